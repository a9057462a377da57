"""Pins the CPU oracle against closed forms, mirroring the reference's own unit tests
(crates/ckks-train-core/src/{modulus,primes,ntt,rns,poly,baseconv}.rs and
crates/ckks-train-ckks/src/{params,keyswitch,evaluator}.rs test modules).

Each check is chosen so that a plausible mistake (a dropped term, a wrong sign or index, a
transposed operand, an FMA-contracted f64 sum) fails it. CPU only.
"""
import math
import random

import numpy as np
import pytest

import oracle as O
from bigint_ref import bitrev, crt, horner, negacyclic, uniform_word, f64_overflow

SEAL_2020 = [
    1152921504598720513, 1099501731841, 1099502518273, 1099502714881, 1099503370241,
    1099503894529, 1099504549889, 1099506515969, 1099507695617, 1099510054913,
    1152921504606584833,
]


# ---------------- modulus.rs:164-216 ----------------

def test_barrett_constant_is_floor_2_128_over_q():
    for q in [2, 3, 4, 1 << 59, (1 << 60) - 1, 1152921504606846883]:
        assert O.barrett(q) == (1 << 128) // q, q


def test_worst_case_mul_and_random():
    rng = random.Random(1)
    for q in [3, (1 << 60) - 1, (1 << 59) + 1, 1152921504606846883, (1 << 50) - 27]:
        m = O.Mod(q)
        a = q - 1
        assert m.mul(a, a) == 1 % q
        assert m.mul(a, 0) == 0
        assert m.mul_shoup(a, a) == 1 % q
        assert m.mul_shoup((1 << 64) - 1, a) == ((1 << 64) - 1) * a % q
        for _ in range(200):
            x, y = rng.randrange(q), rng.randrange(q)
            assert m.mul(x, y) == x * y % q
            z = rng.randrange(1 << 64)
            lazy = m.mul_shoup_lazy(z, y)
            assert lazy < 2 * q and lazy % q == z * y % q


def test_inv_small():
    assert O.Mod(7).inv(3) == 5
    assert O.Mod(7).inv(0) is None
    assert O.Mod(8).inv(4) is None
    assert O.Mod(2).inv(1) == 1


# ---------------- primes.rs:120-191 ----------------

def test_known_primes_and_composites():
    assert O.is_prime(1152921504606846883)
    assert O.is_prime(2) and O.is_prime(3) and O.is_prime((1 << 64) - 59)
    assert not O.is_prime(0) and not O.is_prime(1) and not O.is_prime((1 << 64) - 1)
    for c in [561, 1105, 1729, 3215031751, 3825123056546413051]:
        assert not O.is_prime(c), c
    assert O.is_prime((1 << 61) - 1)
    assert not O.is_prime(4294967291 * 4294967279)


def test_matches_trial_division_below_20000():
    for n in range(20000):
        naive = n >= 2 and all(n % d for d in range(2, int(math.isqrt(n)) + 1))
        assert O.is_prime(n) == naive, n


def test_ntt_prime_search():
    ps = O.ntt_primes(40, 1 << 15, 9)
    for i, q in enumerate(ps):
        assert O.is_prime(q) and (q - 1) % (2 << 15) == 0 and (1 << 39) < q < (1 << 40)
        if i:
            assert q < ps[i - 1]
    # descending and complete: no prime = 1 mod 2n is skipped between consecutive results
    for a, b in zip(ps, ps[1:]):
        for c in range(b + (2 << 15), a, 2 << 15):
            assert not O.is_prime(c)
    ps60 = O.ntt_primes(60, 1 << 15, 3)
    assert O.ntt_primes(60, 1 << 15, 2, skip=1) == ps60[1:]
    assert O.ntt_primes(60, 1 << 15, 0) == []
    assert O.ntt_primes(17, 1 << 15, 1) == [65537]
    with pytest.raises(ValueError):
        O.ntt_primes(17, 1 << 15, 5)
    with pytest.raises(ValueError):
        O.ntt_primes(61, 1 << 15, 1)
    with pytest.raises(ValueError):
        O.ntt_primes(40, 3, 1)
    # the 60-bit prime the reference's NTT spike uses at N = 2^16
    assert O.ntt_primes(60, 1 << 16, 1) == [1152921504606584833]


def test_seal_2020_primes_and_assignment():
    p = O.Params(15, [60] + [40] * 9, [60], 10)
    assert p.q == SEAL_2020[:10] and p.p == SEAL_2020[10:]
    p2 = O.Params(12, [58, 40, 58], [58], 3)
    p58 = O.ntt_primes(58, 1 << 12, 3)
    p40 = O.ntt_primes(40, 1 << 12, 1)
    assert p2.q == [p58[2], p40[0], p58[1]] and p2.p == [p58[0]]


def test_config_params_shape():
    p = O.Params(16, [50] * 24, [60] * 6, 4)
    d50 = O.ntt_primes(50, 1 << 16, 24)
    d60 = O.ntt_primes(60, 1 << 16, 6)
    assert p.q == d50[::-1] and p.p == d60[::-1] and p.alpha == 6


# ---------------- ntt.rs:235-394 ----------------

def test_psi_is_minimal_primitive_root():
    for n, q in [(8, 17), (8, 97), (16, 193), (16, 12289)]:
        expect = next(g for g in range(1, q) if pow(g, n, q) == q - 1)
        assert O.psi(n, q) == expect
    assert O.psi(8, 17) == 3


def test_forward_is_bitreversed_evaluation():
    for n in [16, 256]:
        q = O.ntt_primes(50, n, 1)[0]
        a = O.uniform(n, q, 99, 0)
        f = O.ntt_forward(n, [q], a)
        ps = O.psi(n, q)
        bits = n.bit_length() - 1
        for k in range(n):
            assert int(f[bitrev(k, bits)]) == horner(a, pow(ps, 2 * k + 1, q), q), (n, k)


def test_roundtrip():
    for n, bits in [(2, 20), (4, 30), (256, 40), (1024, 60), (4096, 60), (1 << 13, 50)]:
        q = O.ntt_primes(bits, n, 1)[0]
        a = O.uniform(n, q, n, 1)
        f = O.ntt_forward(n, [q], a)
        assert (f < q).all()
        assert (O.ntt_inverse(n, [q], f) == a).all()


def test_pointwise_product_is_negacyclic_convolution():
    for n, bits in [(8, 60), (64, 60), (128, 50)]:
        q = O.ntt_primes(bits, n, 1)[0]
        a, b = O.uniform(n, q, 1, 0), O.uniform(n, q, 2, 0)
        fa, fb = O.ntt_forward(n, [q], a), O.ntt_forward(n, [q], b)
        fc = np.array([int(x) * int(y) % q for x, y in zip(fa, fb)], dtype=np.uint64)
        assert [int(x) for x in O.ntt_inverse(n, [q], fc)] == negacyclic(a, b, q)


def test_worst_case_and_near_2_60():
    n = 1 << 16
    q = O.ntt_primes(60, n, 1)[0]
    assert q > (1 << 60) - (1 << 40)
    for a in [O.uniform(n, q, 7, 0), np.full(n, q - 1, dtype=np.uint64),
              np.array([q - 1 - (i & 1) for i in range(n)], dtype=np.uint64)]:
        f = O.ntt_forward(n, [q], a)
        assert (f < q).all()
        assert (O.ntt_inverse(n, [q], f) == a).all()
    z = np.zeros(1024, dtype=np.uint64)
    q2 = O.ntt_primes(60, 1024, 1)[0]
    assert (O.ntt_forward(1024, [q2], z) == 0).all()


# ---------------- rns.rs:193-231, poly.rs:637-745 ----------------

def test_galois_permutation_matches_direct_formula():
    for n in [2, 8, 64, 1024]:
        bits = n.bit_length() - 1
        two_n = 2 * n
        gs = [1, 3 % two_n, two_n - 1]
        g = 1
        for _ in range(6):
            g = g * 5 % two_n
            gs.append(g)
        for g in gs:
            perm = O.galois_perm(n, g)
            for i in range(n):
                k = bitrev(i, bits)
                odd = g * (2 * k + 1) % two_n
                assert int(perm[i]) == bitrev((odd - 1) // 2, bits)
            assert sorted(perm.tolist()) == list(range(n))
    assert (O.galois_perm(256, 1) == np.arange(256)).all()


def test_galois_elements():
    n = 1 << 12
    assert O.galois_elt(n, 0) == 1
    assert O.galois_elt(n, 1) == 3
    assert O.galois_elt(n, 5) == pow(3, 5, 2 * n)
    assert O.galois_elt(n, -1) == pow(3, n // 2 - 1, 2 * n)
    assert O.conjugation_elt(n) == 2 * n - 1


def test_automorphism_definition_and_forms():
    n = 256
    qs = O.ntt_primes(50, n, 2)
    a = O.uniform_poly(qs, n, 5)
    for g in [3, 5, 2 * n - 1, O.galois_elt(n, 7)]:
        out = O.automorphism_coeff(n, g, qs, a.ravel()).reshape(2, n)
        for li, q in enumerate(qs):
            want = [0] * n
            for i in range(n):
                e = i * g % (2 * n)
                v = int(a[li, i])
                if e < n:
                    want[e] = v
                else:
                    want[e - n] = (q - v) % q
            assert out[li].tolist() == want
        fa = O.ntt_forward(n, qs, a.ravel())
        lhs = O.automorphism_ntt(n, g, fa)
        rhs = O.ntt_forward(n, qs, out.ravel())
        assert (lhs == rhs).all()


def test_automorphism_is_ring_homomorphism():
    n = 128
    q = O.ntt_primes(50, n, 1)[0]
    a, b = O.uniform(n, q, 11, 0), O.uniform(n, q, 12, 0)
    g = O.galois_elt(n, 3)
    fa, fb = O.ntt_forward(n, [q], a), O.ntt_forward(n, [q], b)
    prod = np.array([int(x) * int(y) % q for x, y in zip(fa, fb)], dtype=np.uint64)
    lhs = O.automorphism_ntt(n, g, prod)
    ra, rb = O.automorphism_ntt(n, g, fa), O.automorphism_ntt(n, g, fb)
    rhs = np.array([int(x) * int(y) % q for x, y in zip(ra, rb)], dtype=np.uint64)
    assert (lhs == rhs).all()


# ---------------- baseconv.rs:610-1240 ----------------

def _conv_ref(x_cols, src, dst, exact):
    out = []
    Q = math.prod(src)
    for t in dst:
        row = []
        for col in x_cols:
            if exact:
                v, _ = crt(col, src)
                row.append(v % t)
            else:
                s = 0
                for j, qj in enumerate(src):
                    qh = Q // qj
                    y = int(col[j]) * pow(qh % qj, -1, qj) % qj
                    s += y * (qh % t)
                row.append(s % t)
        out.append(row)
    return out


def test_convert_matches_seal_formula_and_crt():
    n = 64
    src = O.ntt_primes(50, n, 3)
    dst = O.ntt_primes(60, n, 2) + O.ntt_primes(40, n, 2)
    x = O.uniform_poly(src, n, 21)
    cols = [[int(x[j, c]) for j in range(len(src))] for c in range(n)]
    plain = O.convert(src, dst, x.ravel(), exact=False)
    assert plain.tolist() == _conv_ref(cols, src, dst, False)
    ex = O.convert(src, dst, x.ravel(), exact=True)
    assert ex.tolist() == _conv_ref(cols, src, dst, True)
    # one source prime: plain conversion, already exact
    one = O.convert(src[:1], dst, x[:1].ravel(), exact=True)
    assert one.tolist() == [[int(v) % t for v in x[0]] for t in dst]


def near_tie_residues(primes, n):
    """Coefficient c holds the integer s with s = m or s = D - m (m = c // 2 + 1), D the product
    of `primes`: [x]_D / D lies within n / D of 0 or 1, where the f64 estimate is decided by
    the last bits of the sum."""
    D = math.prod(primes)
    out = []
    for c in range(n):
        m = c // 2 + 1
        s = m if c % 2 == 0 else D - m
        out.append([s % p for p in primes])
    return np.array(out, dtype=np.uint64).T.copy()


def test_convert_exact_f64_sequence_on_near_ties():
    n = 256
    src = O.ntt_primes(60, n, 4)
    dst = O.ntt_primes(50, n, 3)
    x = near_tie_residues(src, n)
    got = O.convert(src, dst, x.ravel(), exact=True)
    plain = O.convert(src, dst, x.ravel(), exact=False)
    Q = math.prod(src)
    diff_fused = 0
    for c in range(n):
        ys = []
        for j, qj in enumerate(src):
            qh = Q // qj
            ys.append(int(x[j, c]) * pow(qh % qj, -1, qj) % qj)
        e = f64_overflow(ys, src)
        diff_fused += e != f64_overflow(ys, src, fused=True)
        for t, p in enumerate(dst):
            assert int(got[t, c]) == (int(plain[t, c]) - e * (Q % p)) % p, (c, t)
    # the inputs really are discriminating: a contracted sum disagrees somewhere
    assert diff_fused > 0


def test_rescale_is_exact_rounding():
    n = 64
    p = O.Params(6, [50, 40, 40, 50], [60], 2)
    for level in [3, 1]:
        x = O.uniform_poly(p.q[: level + 1], n, 31 + level)
        got = p.rescale_by_last(x.ravel(), level, ntt=False)
        ql = p.q[level]
        for c in range(n):
            X, _ = crt(x[:, c], p.q[: level + 1])
            r = (X + (ql - 1) // 2) // ql
            for i in range(level):
                assert int(got[i, c]) == r % p.q[i]
        # NTT form: same values after transforms
        fx = O.ntt_forward(n, p.q[: level + 1], x.ravel())
        gn = p.rescale_by_last(fx, level, ntt=True)
        assert (O.ntt_inverse(n, p.q[:level], gn.ravel()).reshape(level, n) == got).all()


@pytest.mark.parametrize("k", [1, 2, 3])
def test_mod_down_is_exact_rounding(k):
    n = 64
    p = O.Params(6, [50, 50, 40, 40], [60] * k, 2)
    level = 3
    qx = O.uniform_poly(p.q[: level + 1], n, 41)
    px = O.uniform_poly(p.p, n, 42)
    got = p.mod_down(qx.ravel(), level, px.ravel(), ntt=False)
    P = math.prod(p.p)
    for c in range(n):
        X, _ = crt(list(qx[:, c]) + list(px[:, c]), p.q[: level + 1] + p.p)
        r = (X + (P - 1) // 2) // P
        for i in range(level + 1):
            assert int(got[i, c]) == r % p.q[i], (c, i)
    fq = O.ntt_forward(n, p.q[: level + 1], qx.ravel())
    fp = O.ntt_forward(n, p.p, px.ravel())
    gn = p.mod_down(fq, level, fp, ntt=True)
    assert (O.ntt_inverse(n, p.q[: level + 1], gn.ravel()).reshape(level + 1, n) == got).all()


# ---------------- keyswitch.rs:377-684 ----------------

def small_params():
    return O.Params(8, [60, 40, 40, 40, 40, 40], [60, 60], 3)


def test_digit_ranges_clip_to_level():
    p = small_params()
    assert p.digit_ranges(5) == [(0, 2), (2, 4), (4, 6)]
    assert p.digit_ranges(4) == [(0, 2), (2, 4), (4, 5)]
    assert p.digit_ranges(2) == [(0, 2), (2, 3)]
    assert p.digit_ranges(1) == [(0, 2)]
    assert p.digit_ranges(0) == [(0, 1)]
    s = O.Params(10, [60] + [40] * 9, [60], 10)
    assert len(s.digit_ranges(9)) == 10
    assert s.digit_ranges(3) == [(0, 1), (1, 2), (2, 3), (3, 4)]


def _centered_residue(x, D, r):
    return x % r if x <= (D - 1) // 2 else (r - (D - x) % r) % r


@pytest.mark.parametrize("case", ["alpha2_full", "alpha2_clipped", "alpha1", "alpha3"])
def test_exact_lift_matches_crt(case):
    n = 64
    if case.startswith("alpha2"):
        p = O.Params(6, [60, 40, 40, 40, 40, 40], [60, 60], 3)
        level = 5 if case == "alpha2_full" else 4
    elif case == "alpha1":
        p = O.Params(6, [60, 40, 40, 40], [60], 4)
        level = 3
    else:
        p = O.Params(6, [50] * 6, [60, 60], 2)
        level = 4
    x = O.uniform_poly(p.q[: level + 1], n, 13)
    fx = O.ntt_forward(n, p.q[: level + 1], x.ravel())
    digits = p.decompose(fx, level)
    allp = p.q[: level + 1] + p.p
    for (s, e), dig in zip(p.digit_ranges(level), digits):
        coeff = O.ntt_inverse(n, allp, dig.ravel()).reshape(len(allp), n)
        D = math.prod(p.q[s:e])
        for c in range(n):
            X, _ = crt(x[s:e, c], p.q[s:e])
            for li, r in enumerate(allp):
                assert int(coeff[li, c]) == _centered_residue(X, D, r), (case, li, c)
        for i in range(s, e):
            assert (dig[i] == fx.reshape(level + 1, n)[i]).all()


def _ip_reference(p, digits, level, key, kml):
    """Fully reduced inner product (every product and sum reduced), then mod_down."""
    parts, lp = p.key_shape(kml)
    key = key.reshape(parts, 2, lp, p.n)
    limbs = list(range(level + 1)) + [kml + 1 + t for t in range(p.k)]
    primes = p.q[: level + 1] + p.p
    acc = np.zeros((2, len(limbs), p.n), dtype=object)
    for j in range(len(digits)):
        for li, (kl, m) in enumerate(zip(limbs, primes)):
            d = digits[j, li].astype(object)
            for s in range(2):
                acc[s, li] = (acc[s, li] + d * key[j, s, kl].astype(object)) % m
    acc = acc.astype(np.uint64)
    kb = p.mod_down(acc[0, : level + 1].ravel(), level, acc[0, level + 1:].ravel())
    ka = p.mod_down(acc[1, : level + 1].ravel(), level, acc[1, level + 1:].ravel())
    return kb, ka


def test_apply_key_matches_reduced_reference_and_extremes():
    p = O.Params(7, [60] + [40] * 5, [60], 6)  # alpha = 1, one special prime
    p2 = O.Params(7, [50] * 6, [60, 60], 3)
    for pp in (p, p2):
        level = len(pp.q) - 1
        key = pp.uniform_key(level, 5)
        x = O.uniform_poly(pp.q, pp.n, 17)
        fx = O.ntt_forward(pp.n, pp.q, x.ravel())
        digits = pp.decompose(fx, level)
        worst = digits.copy()
        allp = pp.q + pp.p
        for li, m in enumerate(allp):
            worst[:, li, :] = m - 1
        for d in (digits, worst):
            got = pp.apply_key(d, level, key, level)
            want = _ip_reference(pp, d, level, key, level)
            assert (got[0] == want[0]).all() and (got[1] == want[1]).all()


def test_key_switch_is_apply_key_of_decompose_and_capped_key():
    p = O.Params(8, [50] * 6, [60, 60], 3)
    key = p.uniform_key(5, 3)
    level = 3
    x = O.ntt_forward(p.n, p.q[: level + 1], O.uniform_poly(p.q[: level + 1], p.n, 4).ravel())
    kb, ka = p.key_switch(x, level, key, 5)
    kb2, ka2 = p.apply_key(p.decompose(x, level), level, key, 5)
    assert (kb == kb2).all() and (ka == ka2).all()
    # a key uploaded only for levels <= 3 gives the same switch (capped keys)
    parts, lp = p.key_shape(5)
    full = key.reshape(parts, 2, lp, p.n)
    cparts, clp = p.key_shape(3)
    capped = np.concatenate([full[:cparts, :, :4], full[:cparts, :, 6:]], axis=2)
    kb3, ka3 = p.key_switch(x, level, capped, 3)
    assert (kb == kb3).all() and (ka == ka3).all()


def test_hoisted_matches_direct():
    for pp, level in [(O.Params(8, [60, 40, 40, 40], [60], 4), 3),
                      (O.Params(8, [60, 40, 40, 40, 40, 40], [60, 60], 3), 5),
                      (O.Params(8, [60, 40, 40, 40, 40, 40], [60, 60], 3), 4)]:
        kml = len(pp.q) - 1
        c0 = O.ntt_forward(pp.n, pp.q[: level + 1], O.uniform_poly(pp.q[: level + 1], pp.n, 50).ravel())
        c1 = O.ntt_forward(pp.n, pp.q[: level + 1], O.uniform_poly(pp.q[: level + 1], pp.n, 51).ravel())
        elts = [O.galois_elt(pp.n, 5), O.conjugation_elt(pp.n), O.galois_elt(pp.n, 1)]
        keys = [pp.uniform_key(kml, 60 + i) for i in range(len(elts))]
        outs = pp.hoisted_galois(c0, c1, level, elts, keys, kml)
        for s, (g, key) in enumerate(zip(elts, keys)):
            o0, o1 = pp.apply_galois(c0, c1, level, g, key, kml)
            assert (outs[s, 0] == o0).all() and (outs[s, 1] == o1).all()


def test_relinearize_rescale_mul_compositions():
    p = O.Params(8, [50] * 6, [60, 60], 3)
    level = 5
    key = p.uniform_key(level, 7)
    c = [O.ntt_forward(p.n, p.q, O.uniform_poly(p.q, p.n, 70 + i).ravel()).reshape(6, p.n) for i in range(4)]
    d0, d1, d2 = p.mul_no_relin(c[0], c[1], c[2], c[3], level)
    for i, q in enumerate(p.q):
        a0, a1, b0, b1 = (x[i].astype(object) for x in c)
        assert d0[i].tolist() == list(a0 * b0 % q)
        assert d1[i].tolist() == list((a0 * b1 + a1 * b0) % q)
        assert d2[i].tolist() == list(a1 * b1 % q)
    o0, o1 = p.relinearize(d0, d1, d2, level, key, level)
    kb, ka = p.key_switch(d2, level, key, level)
    for i, q in enumerate(p.q):
        assert o0[i].tolist() == list((d0[i].astype(object) + kb[i].astype(object)) % q)
        assert o1[i].tolist() == list((d1[i].astype(object) + ka[i].astype(object)) % q)
    r = p.rescale_ct(np.stack([o0, o1]), level)
    assert (r[0] == p.rescale_by_last(o0, level)).all()
    assert (r[1] == p.rescale_by_last(o1, level)).all()


# ---------------- seeded inputs ----------------

def test_uniform_generator_definition():
    for q in [17, (1 << 50) - 27, (1 << 60) - 93]:
        got = O.uniform(64, q, 123, 9)
        assert got.tolist() == [uniform_word(123, 9, i, q) for i in range(64)]
    big = O.uniform(1 << 16, (1 << 60) - 93, 5, 0)
    assert (big < (1 << 60) - 93).all()
    # roughly uniform: mean within 1% of q/2
    assert abs(float(big.astype(np.float64).mean()) / ((1 << 60) / 2) - 1) < 0.01


# ---------------- scalar products and exact labels (evaluator.rs:330-418,738-797) ----------------

def test_poly_mul_scalar_is_the_big_integer_product():
    p = O.Params(8, [60, 50, 40], [60], 3)
    level = 2
    a = O.uniform_poly(p.q, p.n, 81)
    a[0, :3] = [0, 1, p.q[0] - 1]
    for k in [0, 1, 3, (1 << 62) - 1, p.q[1], p.q[1] + 5, 123456789123456789]:
        for neg in (False, True):
            got = p.poly_mul_scalar(a.ravel(), level, k, neg)
            for i, q in enumerate(p.q):
                want = [(int(x) * k) % q for x in a[i]]
                if neg:
                    want = [(q - w) % q for w in want]
                assert got[i].tolist() == want, (k, neg, i)


def _ulp_neighbours(x, k):
    import struct
    b = struct.unpack("<Q", struct.pack("<d", x))[0]
    return [struct.unpack("<d", struct.pack("<Q", b + k))[0], struct.unpack("<d", struct.pack("<Q", b - k))[0]]


def test_exact_product_scale_divides_to_the_target_and_is_the_nearest_such_label():
    p = O.Params(10, [60, 40, 40, 40, 40, 40], [60, 60], 3)
    rng = random.Random(5)
    delta = 2.0 ** 40
    targets = [delta, 1.5 * delta, delta * delta / p.q[4], 3.0, 2.0 ** 70] + \
              [rng.uniform(2.0 ** 20, 2.0 ** 60) for _ in range(200)]
    for t in targets:
        for level in range(1, len(p.q)):
            q = float(p.q[level])
            try:
                lab = p.exact_product_scale(level, t)
            except ValueError as e:
                # then no candidate within 8 ulps of t * q divides back to t
                assert str(e) == "ScaleMismatch"
                assert all(c / q != t for k in range(9) for c in _ulp_neighbours(t * q, k))
                continue
            assert lab / q == t
            # no candidate nearer to t * q (in ulps) divides back to t
            k = next(k for k in range(9) if lab in _ulp_neighbours(t * q, k))
            assert all(c / q != t for j in range(k) for c in _ulp_neighbours(t * q, j))
            if lab != _ulp_neighbours(t * q, k)[0]:
                assert _ulp_neighbours(t * q, k)[0] / q != t  # above is tried before below


def test_exact_product_scale_errors():
    p = O.Params(10, [60, 40, 40], [60], 3)
    for args, kind in [((3, 2.0 ** 40), "InvalidParams"), ((0, 2.0 ** 40), "NoLevelsLeft"),
                       ((1, 0.0), "InvalidParams"), ((1, -1.0), "InvalidParams"),
                       ((1, float("nan")), "InvalidParams"), ((1, float("inf")), "InvalidParams"),
                       ((1, 1e300), "InvalidParams")]:
        with pytest.raises(ValueError, match=kind):
            p.exact_product_scale(*args)


def test_scalar_rounding_is_half_away_from_zero():
    from decimal import Decimal, ROUND_HALF_UP
    p = O.Params(10, [60, 40, 40], [60], 3)
    lab = p.exact_product_scale(2, 2.0 ** 40)
    src = lab / 2.0 ** 30  # plaintext scale 2^30
    for value in [0.5, -0.5, 1.5, -2.5, 1e-3, -1.25, 3.0, 1.0 / 3.0, 2.0 ** -30 * 2.5]:
        label, sp, k = p.mul_scalar_parts(2, src, value, 2.0 ** 40)
        x = value * sp
        want = float(Decimal(x).quantize(Decimal(1), rounding=ROUND_HALF_UP))
        assert label == lab and k == want, value
    # the reference's bound: an integer at or above 2^62 is rejected (evaluator.rs:1754-1758 shape)
    with pytest.raises(ValueError, match="InvalidParams"):
        p.mul_scalar_parts(2, 2.0 ** 40, 1e9, 2.0 ** 40 * 1e9)
    with pytest.raises(ValueError, match="InvalidParams"):
        p.mul_scalar_parts(2, 2.0 ** 40, float("nan"), 2.0 ** 40)
    with pytest.raises(ValueError, match="InvalidParams"):  # plaintext scale below 2^20
        p.mul_scalar_parts(2, 2.0 ** 40 * p.q[2], 1.0, 2.0 ** 40)


# ---------------- raise from level 0 (poly.rs:298-345, 637-745 tests) ----------------

def test_raise_from_level0_is_the_centered_lift():
    for qb in ([50, 40, 40, 30], [60, 50, 50, 40]):
        p = O.Params(6, qb, [60], 2)
        q0 = p.q[0]
        a = O.uniform_poly(p.q[:1], p.n, 17)[0]
        a[:6] = [0, 1, q0 // 2, q0 // 2 + 1, q0 - 1, q0 - 2]
        top = len(p.q) - 1
        coeff = p.raise_from_level0(a, top, ntt=False)
        for i, q in enumerate(p.q):
            for c in range(p.n):
                x = int(a[c])
                centered = x if x <= (q0 - 1) // 2 else x - q0
                assert int(coeff[i, c]) == centered % q, (i, c)
        # NTT form: limb 0 untouched, the rest = forward(centered lift of inverse(limb 0))
        a_ntt = O.ntt_forward(p.n, p.q[:1], a.copy())
        got = p.raise_from_level0(a_ntt, top, ntt=True)
        assert (got[0] == a_ntt).all()
        want = O.ntt_forward(p.n, p.q[: top + 1], coeff.ravel().copy()).reshape(top + 1, p.n)
        assert (got[1:] == want[1:]).all()
        assert (p.raise_from_level0(a, 0, ntt=False) == a).all()


def test_params_ntt_matches_the_standalone_transform():
    p = O.Params(9, [50, 40, 60], [60], 1)
    a = O.uniform_poly(p.q, p.n, 5)
    f = p.ntt(a, 3, True)
    assert (f.ravel() == O.ntt_forward(p.n, p.q, a.ravel().copy())).all()
    assert (p.ntt(f, 3, False) == a).all()


# ---------------- bootstrap/linear.rs:921-945 (mul_by_x_half) ----------------

def test_mul_by_x_half_is_the_monomial_product():
    """X^(N/2) a: equal to the schoolbook negacyclic product with the monomial (coefficient
    form), to the NTT-form product with NTT(X^(N/2)) (the reference's own method), applying it
    twice negates (X^N = -1), and every NTT slot is multiplied by a square root of -1."""
    p = O.Params(5, [50, 40], [60], 1)
    level = 1
    a_coef = O.uniform_poly(p.q, p.n, 31)
    a = p.ntt(a_coef, 2, True)
    got = p.mul_by_x_half(a, level)
    mono = np.zeros((2, p.n), dtype=np.uint64)
    mono[:, p.n // 2] = 1
    for i, q in enumerate(p.q):
        want = negacyclic([int(x) for x in a_coef[i]], [int(x) for x in mono[i]], q)
        assert [int(x) for x in p.ntt(got, 2, False)[i]] == want
    assert (got == p.poly_mul(a, p.ntt(mono, 2, True), level)).all()
    neg = p.poly_sub(np.zeros_like(a), a, level)
    assert (p.mul_by_x_half(got, level) == neg).all()
    for i, q in enumerate(p.q):
        for x, y in zip(a[i], got[i]):
            r = int(y) * pow(int(x), -1, q) % q
            assert r * r % q == q - 1


def test_poly_sub_wraps():
    p = O.Params(4, [30, 30], [40], 1)
    a = np.zeros((2, p.n), dtype=np.uint64)
    b = np.ones((2, p.n), dtype=np.uint64)
    assert (p.poly_sub(a, b, 1) == np.array(p.q, dtype=np.uint64)[:, None] - 1).all()
    assert (p.poly_sub(b, a, 1) == 1).all()
