"""Device parity: every HIP stage against the CPU oracle on identical seeded inputs, bit for
bit, through the C ABI (mirrors the reference's GPU case suite,
crates/ckks-train-cuda/tests/cases/{mod,ks}.rs: seeded uniform residues, all-(q-1) extremes,
levels spread so clipped digits occur, near-tie inputs for the f64 overflow estimate)."""
import sys

import numpy as np
import pytest

import ckh
import oracle as O
from gpu_cases import (ALPHA1, C2, C2_SHAPE_N4096, C2_SHAPE_SMALL_N, C4, C4_SHAPE_SMALL_N, SMALL, SMALL_N8192,
                       c4_entries_mask, host_uniform, levels_for, make, near_tie_digits, near_tie_p_part,
                       oracle_matvec)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=[SMALL, ALPHA1, C2_SHAPE_SMALL_N, C4_SHAPE_SMALL_N, C2_SHAPE_N4096, SMALL_N8192],
                ids=lambda c: c[0])
def env(request):
    say("case", request.param[0])
    ctx, ref = make(request.param)
    yield ctx, ref
    ctx.synchronize()


# ---------------------------------------------------------------- NTT

@pytest.mark.parametrize("log_n,bits,limbs", [(1, 20, 1), (4, 30, 2), (10, 60, 3), (12, 50, 2), (13, 50, 4),
                                              (14, 60, 2), (15, 50, 3), (16, 50, 3), (16, 60, 2), (17, 60, 1)])
def test_ntt_matches_oracle(log_n, bits, limbs):
    say("ntt", log_n, bits, limbs)
    n = 1 << log_n
    q = O.ntt_primes(bits, n, limbs)
    p = O.ntt_primes(bits, n, limbs + 1)[limbs:]
    ctx = ckh.Context(log_n, q, p, 1)
    a = O.uniform_poly(q, n, 7 + log_n)
    d = ctx.upload(a, limbs - 1, ntt=False)
    f = d.ntt()
    want = O.ntt_forward(n, q, a.ravel()).reshape(limbs, n)
    assert (f.download() == want).all()
    assert (f.intt().download() == a).all()
    # extremes: all q - 1 and zero
    mx = np.stack([np.full(n, qi - 1, dtype=np.uint64) for qi in q])
    fm = ctx.upload(mx, limbs - 1, ntt=False).ntt()
    assert (fm.download() == O.ntt_forward(n, q, mx.ravel()).reshape(limbs, n)).all()
    assert (fm.intt().download() == mx).all()
    z = ctx.upload(np.zeros((limbs, n), dtype=np.uint64), limbs - 1, ntt=False).ntt()
    assert (z.download() == 0).all()


@pytest.mark.parametrize("log_n", [12, 13, 14, 15, 16])
def test_one_launch_forward_ntt(log_n):
    """The one-launch transforms (the first pass handed to the second through one XCD's L2),
    forward and inverse, against the oracle and against the two-launch forms, for mixed 50/60-bit limbs in limb counts
    that leave XCDs with no limb, one, or uneven shares, in place and out of place, at several
    lags between a limb's column and row tiles; the device error word stays clear."""
    n = 1 << log_n
    q = O.ntt_primes(50, n, 5)
    p = O.ntt_primes(60, n, 3)
    ctx = ckh.Context(log_n, q, p, 1)
    assert ctx.tuning("xcds") in (0, 8)
    if ctx.tuning("xcds") == 0:
        pytest.skip("XCD layout not verified on this device: the one-launch transform is off")
    for lag in (1, 4, 16):
        ctx.set_tuning("fused_lag", lag)
        for level, pl in ((0, 0), (2, 3), (4, 3)):
            say("one-launch ntt", log_n, "lag", lag, "limbs", level + 1 + pl)
            primes = q[: level + 1] + (p if pl else [])
            a = O.uniform_poly(primes, n, 31 + lag + level)
            want = O.ntt_forward(n, primes, a.ravel()).reshape(len(primes), n)
            d = ctx.upload(a, level, p_limbs=pl, ntt=False)
            f = ctx.upload(want, level, p_limbs=pl, ntt=True)
            ctx.set_tuning("ntt_fused", 1)
            got, inv = d.ntt().download(), f.intt().download()
            ctx.set_tuning("ntt_fused", 0)
            two, inv2 = d.ntt().download(), f.intt().download()
            ctx.set_tuning("ntt_fused", 1)
            assert (got == want).all() and (two == want).all(), (lag, level, pl)
            assert (inv == a).all() and (inv2 == a).all(), (lag, level, pl)
    # many limbs in one launch (every XCD busy with several), in place through the timing hook
    polys = [ctx.uniform(4, seed=70 + i, p_limbs=3, ntt=False) for i in range(13)]
    hosts = [x.download() for x in polys]
    ctx.time_ntt(polys, forward=True, reps=1)  # two in-place transforms of every limb
    primes = q + p
    for h, x in zip(hosts, polys):
        once = O.ntt_forward(n, primes, h.ravel())
        # the transform applied twice to the same words (the warm-up launch and the timed one)
        twice = O.ntt_forward(n, primes, once).reshape(len(primes), n)
        assert (x.download() == twice).all()
    ctx.synchronize()  # raises if a wait overran its bound


def test_c1_plain_ntt_round_trip():
    """N = 2^13, the 4 largest 50-bit primes = 1 mod 2N, forward then inverse, bit-exact."""
    log_n, n = 13, 1 << 13
    q = O.ntt_primes(50, n, 4)
    p = O.ntt_primes(60, n, 1)
    ctx = ckh.Context(log_n, q, p, 1)
    x = ctx.uniform(3, seed=2024, ntt=False)
    a = O.uniform_poly(q, n, 2024)
    assert (x.download() == a).all()
    f = x.ntt()
    assert (f.download() == O.ntt_forward(n, q, a.ravel()).reshape(4, n)).all()
    assert (f.intt().download() == a).all()


# ---------------------------------------------------------------- pointwise stages

def test_uniform_generator_matches_oracle(env):
    ctx, ref = env
    level = len(ref.q) - 1
    x = ctx.uniform(level, seed=99, stream0=5, p_limbs=ref.k)
    assert (x.download() == host_uniform(ref, level, 99, 5, p_limbs=1)).all()


def test_automorphism(env):
    ctx, ref = env
    level = len(ref.q) - 1
    a = host_uniform(ref, level, 3)
    d = ctx.upload(a, level)
    for g in [ckh.galois_elt(ref.log_n, 1), ckh.galois_elt(ref.log_n, -3), ckh.conjugation_elt(ref.log_n)]:
        got = d.automorphism(g).download()
        assert (got == O.automorphism_ntt(ref.n, g, a.ravel()).reshape(a.shape)).all()


def test_automorphism_coefficient_form(env):
    """The signed permutation of coefficient-form input (poly.rs:255-276) against the oracle, on a
    Q polynomial and a Q*P one (each limb negated mod its own prime), zero words included; it
    commutes with the NTT-form permutation through the transform."""
    ctx, ref = env
    level = len(ref.q) - 1
    a = host_uniform(ref, level, 5)
    a[0, :7] = 0
    qp = host_uniform(ref, level, 6, p_limbs=1)
    steps = [1, -3, 7]
    for g in [ckh.galois_elt(ref.log_n, s) for s in steps] + [ckh.conjugation_elt(ref.log_n), 1]:
        got = ctx.upload(a, level, ntt=False).automorphism(g).download()
        assert (got == O.automorphism_coeff(ref.n, g, ref.q[: level + 1], a.ravel()).reshape(a.shape)).all(), g
        got = ctx.upload(qp, level, p_limbs=len(ref.p), ntt=False).automorphism(g).download()
        want = O.automorphism_coeff(ref.n, g, ref.q[: level + 1] + ref.p, qp.ravel()).reshape(qp.shape)
        assert (got == want).all(), g
    g = ckh.galois_elt(ref.log_n, 5)
    c = ctx.upload(a, level, ntt=False)
    via_coeff = c.automorphism(g).ntt().download()
    via_ntt = c.ntt().automorphism(g).download()
    assert (via_coeff == via_ntt).all()
    with pytest.raises(ckh.InvalidParams):
        c.automorphism(2)


def test_rescale_poly(env):
    ctx, ref = env
    for level in levels_for(len(ref.q)):
        if level == 0:
            with pytest.raises(ckh.NoLevelsLeft):
                ctx.upload(host_uniform(ref, 0, 1), 0).rescale_by_last()
            continue
        a = host_uniform(ref, level, 11 + level)
        got = ctx.upload(a, level).rescale_by_last().download()
        assert (got == ref.rescale_by_last(a.ravel(), level)).all(), level


def test_mod_down_random_and_near_ties(env):
    ctx, ref = env
    for level in levels_for(len(ref.q)):
        qpart = host_uniform(ref, level, 21 + level)
        for ppart in (O.uniform_poly(ref.p, ref.n, 31 + level), near_tie_p_part(ref)):
            qp = np.concatenate([qpart, ppart])
            got = ctx.upload(qp, level, p_limbs=ref.k).mod_down().download()
            want = ref.mod_down(qpart.ravel(), level, ppart.ravel())
            assert (got == want).all(), level


# ---------------------------------------------------------------- key switching

def test_decompose_random_extreme_near_tie(env):
    ctx, ref = env
    for level in levels_for(len(ref.q)):
        inputs = [host_uniform(ref, level, 41 + level),
                  np.stack([np.full(ref.n, q - 1, dtype=np.uint64) for q in ref.q[: level + 1]])]
        if level >= 1:
            inputs.append(near_tie_digits(ref, level))
        for a in inputs:
            digs = ctx.decompose(ctx.upload(a, level))
            want = ref.decompose(a.ravel(), level)
            assert len(digs) == len(want)
            for j, d in enumerate(digs):
                assert (d.download() == want[j]).all(), (level, j)


def test_key_switch_levels_and_capped_key(env):
    ctx, ref = env
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=5)
    hkey = ref.uniform_key(L, 5)
    for level in levels_for(len(ref.q)):
        a = host_uniform(ref, level, 51 + level)
        kb, ka = ctx.key_switch(ctx.upload(a, level), key)
        wb, wa = ref.key_switch(a.ravel(), level, hkey, L)
        assert (kb.download() == wb).all() and (ka.download() == wa).all(), level
    # a key uploaded for levels <= L // 2 switches those levels identically
    cap = L // 2
    ckey = ctx.ksk_uniform(cap, seed=6)
    hcap = ref.uniform_key(cap, 6)
    a = host_uniform(ref, cap, 77)
    kb, ka = ctx.key_switch(ctx.upload(a, cap), ckey)
    wb, wa = ref.key_switch(a.ravel(), cap, hcap, cap)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()
    with pytest.raises(ckh.InvalidParams):
        ctx.key_switch(ctx.upload(host_uniform(ref, cap + 1, 78), cap + 1), ckey)


def test_key_switch_worst_case_inputs(env):
    ctx, ref = env
    L = len(ref.q) - 1
    parts, lp = ref.key_shape(L)
    primes = ref.q + ref.p
    hkey = np.stack([np.stack([np.full(ref.n, q - 1, dtype=np.uint64) for q in primes])
                     for _ in range(2 * parts)])
    key = ctx.ksk_upload(hkey, L)
    a = np.stack([np.full(ref.n, q - 1, dtype=np.uint64) for q in ref.q])
    kb, ka = ctx.key_switch(ctx.upload(a, L), key)
    wb, wa = ref.key_switch(a.ravel(), L, hkey, L)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()


def test_decompose_apply_equals_key_switch(env):
    ctx, ref = env
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=8)
    a = ctx.uniform(L, seed=61)
    digs = ctx.decompose(a)
    kb, ka = ctx.apply_key(digs, key)
    kb2, ka2 = ctx.key_switch(a, key)
    assert (kb.download() == kb2.download()).all() and (ka.download() == ka2.download()).all()


def test_batch_equals_individual_switches(env):
    ctx, ref = env
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=9)
    rng = np.random.default_rng(3)
    lv = [int(x) for x in rng.integers(0, L + 1, size=9)]
    polys = [ctx.uniform(l, seed=100 + i) for i, l in enumerate(lv)]
    batch = ctx.key_switch_batch(polys, key)
    for p, (kb, ka) in zip(polys, batch):
        sb, sa = ctx.key_switch(p, key)
        assert (kb.download() == sb.download()).all() and (ka.download() == sa.download()).all()


# ---------------------------------------------------------------- evaluator (HeOps)

def _ct(ctx, ref, level, seed, ncomp=2):
    polys = np.stack([host_uniform(ref, level, seed + i) for i in range(ncomp)])
    return ctx.ct_upload(polys, level, 2.0 ** 40), polys


def test_relinearize_rotate_conjugate_rescale(env):
    ctx, ref = env
    L = len(ref.q) - 1
    n = ref.n
    relin = ctx.ksk_uniform(L, seed=12)
    steps = [1, 5, -2]
    elts = [ckh.galois_elt(ref.log_n, s) for s in steps] + [ckh.conjugation_elt(ref.log_n)]
    gkeys = [ctx.ksk_uniform(L, seed=20 + i, galois=g) for i, g in enumerate(elts)]
    ev = ckh.Evaluator(ctx, relin, gkeys)
    for level in (L, L // 2, 0):
        ct3, h3 = _ct(ctx, ref, level, 300 + level, ncomp=3)
        r = ev.relinearize(ct3)
        w0, w1 = ref.relinearize(h3[0], h3[1], h3[2], level, ref.uniform_key(L, 12), L)
        got = r.download()
        assert (got[0] == w0).all() and (got[1] == w1).all(), level
        assert r.scale == 2.0 ** 40 and r.level == level
        ct2, h2 = _ct(ctx, ref, level, 400 + level)
        for i, s in enumerate(steps):
            rr = ev.rotate(ct2, s).download()
            w0, w1 = ref.apply_galois(h2[0], h2[1], level, elts[i], ref.uniform_key(L, 20 + i), L)
            assert (rr[0] == w0).all() and (rr[1] == w1).all(), (level, s)
        cc = ev.conjugate(ct2).download()
        w0, w1 = ref.apply_galois(h2[0], h2[1], level, elts[-1], ref.uniform_key(L, 23), L)
        assert (cc[0] == w0).all() and (cc[1] == w1).all()
        if level > 0:
            rs = ev.rescale(ct2)
            assert (rs.download() == ref.rescale_ct(h2.ravel(), level)).all()
            assert rs.scale == 2.0 ** 40 / float(ref.q[level])
    c = ev.counter()
    assert c.relins == 3 and c.rotations == 12 and c.rescales == 2
    assert c.key_switches() == 15


def test_hoisted_rotate_matches_oracle_and_direct(env):
    ctx, ref = env
    L = len(ref.q) - 1
    steps = [1, 2, 3, 7, -1, 0, ref.n // 2]
    nontriv = [s for s in steps if s % (ref.n // 2)]
    elts = {s: ckh.galois_elt(ref.log_n, s) for s in nontriv}
    gkeys = [ctx.ksk_uniform(L, seed=50 + i, galois=elts[s]) for i, s in enumerate(nontriv)]
    ev = ckh.Evaluator(ctx, None, gkeys)
    for level in (L, L // 2):
        ct, h = _ct(ctx, ref, level, 500 + level)
        outs = ev.hoisted_rotate(ct, steps)
        want = ref.hoisted_galois(h[0], h[1], level, [elts[s] for s in nontriv],
                                  [ref.uniform_key(L, 50 + i) for i in range(len(nontriv))], L)
        k = 0
        for s, o in zip(steps, outs):
            got = o.download()
            if s in elts:
                assert (got[0] == want[k, 0]).all() and (got[1] == want[k, 1]).all(), (level, s)
                direct = ev.rotate(ct, s).download()
                assert (got == direct).all()
                k += 1
            else:
                assert (got == h).all()


def test_mul_add_sub_negate(env):
    ctx, ref = env
    L = len(ref.q) - 1
    relin = ctx.ksk_uniform(L, seed=13)
    ev = ckh.Evaluator(ctx, relin)
    level = max(L // 2, 1)
    a, ha = _ct(ctx, ref, level, 600)
    b, hb = _ct(ctx, ref, level, 610)
    d = ev.mul_no_relin(a, b).download()
    w = ref.mul_no_relin(ha[0], ha[1], hb[0], hb[1], level)
    for i in range(3):
        assert (d[i] == w[i]).all()
    m = ev.mul(a, b).download()
    w0, w1 = ref.relinearize(w[0], w[1], w[2], level, ref.uniform_key(L, 13), L)
    assert (m[0] == w0).all() and (m[1] == w1).all()
    s = ev.add(a, b).download()
    df = ev.sub(a, b).download()
    ng = ev.negate(a).download()
    for i, q in enumerate(ref.q[: level + 1]):
        for c in range(2):
            x, y = ha[c, i].astype(object), hb[c, i].astype(object)
            assert s[c, i].tolist() == list((x + y) % q)
            assert df[c, i].tolist() == list((x - y) % q)
            assert ng[c, i].tolist() == list((-x) % q)
    assert ev.counter().ct_ct_tensor == 2 and ev.counter().relins == 1


def test_evaluator_errors(env):
    ctx, ref = env
    L = len(ref.q) - 1
    ev = ckh.Evaluator(ctx)
    a, _ = _ct(ctx, ref, L, 700)
    b, _ = _ct(ctx, ref, max(L - 1, 0), 710)
    if L > 0:
        with pytest.raises(ckh.LevelMismatch):
            ev.add(a, b)
    with pytest.raises(ckh.InvalidParams, match="relinearization key"):
        ev.relinearize(_ct(ctx, ref, L, 720, ncomp=3)[0])
    with pytest.raises(ckh.InvalidParams, match="missing galois key"):
        ev.rotate(a, 1)
    with pytest.raises(ckh.InvalidParams):
        ev.mul_no_relin(_ct(ctx, ref, L, 730, ncomp=3)[0], a)
    z, _ = _ct(ctx, ref, 0, 740)
    with pytest.raises(ckh.NoLevelsLeft):
        ev.rescale(z)
    with pytest.raises(ckh.NoLevelsLeft):
        ev.mul_no_relin(z, z)
    c2 = ctx.ct_upload(np.stack([host_uniform(ref, L, 1), host_uniform(ref, L, 2)]), L, 2.0 ** 30)
    with pytest.raises(ckh.ScaleMismatch):
        ev.add(a, c2)
    assert ev.rotate(a, ref.n // 2).download().tolist() == a.download().tolist()
    assert ev.counter().key_switches() == 0


# ---------------------------------------------------------------- full size (C2)

def test_c2_full_size_key_switch():
    """N = 2^16, 24 x 50-bit q, 6 x 60-bit p, dnum 4, level 23: bit-exact against the oracle."""
    say("c2 full size")
    ctx, ref = make(C2)
    L = 23
    key = ctx.ksk_uniform(L, seed=424242)
    d = ctx.uniform(L, seed=4242)
    kb, ka = ctx.key_switch(d, key)
    O.set_threads(16)
    wb, wa = ref.key_switch(host_uniform(ref, L, 4242).ravel(), L, ref.uniform_key(L, 424242), L)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()


def test_c3_full_batch():
    """The benchmark workload at its stated size: 256 polynomials at N = 2^16 (24 + 6 primes,
    dnum 4), limbs U{4..24} from its seed, one shared key. The batched switch, on its first run
    and on graph replays, equals one switch per call for every member (by digest), and eight
    members whose levels span the workload's range equal the oracle word for word."""
    say("c3 full batch")
    ctx, ref = make(C2)
    rng = np.random.default_rng(20261202)
    levels = [int(x) - 1 for x in rng.integers(4, 25, size=256)]
    key = ctx.ksk_uniform(23, seed=515151)
    polys = [ctx.uniform(lv, seed=5150 + i) for i, lv in enumerate(levels)]
    single = [(kb.digest(), ka.digest()) for kb, ka in (ctx.key_switch(p, key) for p in polys)]
    for rep in range(3):  # direct, captured, replayed
        got = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
        assert got == single, rep
    batch = ctx.key_switch_batch(polys, key)
    distinct = sorted(set(levels))
    picks = [distinct[round(j * (len(distinct) - 1) / 7)] for j in range(8)]
    O.set_threads(16)
    wkey = ref.uniform_key(23, 515151)
    for lv in picks:
        i = levels.index(lv)
        say("oracle member", i, "level", lv)
        wb, wa = ref.key_switch(host_uniform(ref, lv, 5150 + i).ravel(), lv, wkey, 23)
        kb, ka = batch[i]
        assert (kb.download() == wb).all() and (ka.download() == wa).all(), (i, lv)
    assert min(picks) == min(levels) and max(picks) == max(levels)


# ---------------------------------------------------------------- BSGS (hoisted rotations)

def _bsgs_case(case, baby, giant, mask, seed, level=None):
    ctx, ref = make(case)
    L = len(ref.q) - 1
    level = L if level is None else level
    steps = sorted({k2 for k2 in range(1, baby)} | {baby * k1 for k1 in range(1, giant)})
    keys, hkeys = [], {}
    for i, st in enumerate(steps):
        g = ckh.galois_elt(ref.log_n, st)
        keys.append(ctx.ksk_uniform(L, seed=seed + 10 * i, galois=g))
        hkeys[st] = seed + 10 * i
    ev = ckh.Evaluator(ctx, None, keys)
    c0, c1 = host_uniform(ref, level, seed + 1), host_uniform(ref, level, seed + 2)
    x = ctx.ct_upload(np.stack([c0, c1]), level, 2.0 ** 40)
    entries_h, entries_d = [], []
    for k1 in range(giant):
        for k2 in range(baby):
            if mask[k1][k2]:
                w = host_uniform(ref, level, seed + 1000 + k1 * baby + k2)
                entries_h.append(w)
                entries_d.append(ctx.upload(w, level))
            else:
                entries_h.append(None)
                entries_d.append(None)
    got = ev.matvec_plain(x, entries_d, 2.0 ** 40, baby, giant)
    key_cache = {}

    def key_of_step(st):
        if st not in key_cache:
            key_cache[st] = ref.uniform_key(L, hkeys[st])
        return key_cache[st]

    want = oracle_matvec(ref, c0, c1, level, entries_h, baby, giant, 1, key_of_step, L)
    g = got.download()
    assert (g[0] == want[0]).all() and (g[1] == want[1]).all()
    assert got.scale == 2.0 ** 80
    nrot = sum(1 for k2 in range(1, baby) if any(mask[k1][k2] for k1 in range(giant)))
    nrot += sum(1 for k1 in range(1, giant) if any(mask[k1]))
    c = ev.counter()
    assert c.rotations == nrot and c.ct_pt_mults == sum(map(sum, mask))


def test_bsgs_matvec_small():
    mask = [[True, False, True, True], [False, False, False, False], [True, True, False, False],
            [False, True, True, True]]
    _bsgs_case(C4_SHAPE_SMALL_N, 4, 4, mask, seed=900)


def test_c4_hoisted_bsgs_full_size():
    """N = 2^15, 20 x 50-bit q, 4 x 60-bit p, dnum 5: 32 hoisted baby steps, 32 giant steps,
    64 diagonals (two per giant block), against the oracle composition."""
    say("c4 full size")
    O.set_threads(16)
    _bsgs_case(C4, 32, 32, c4_entries_mask(), seed=4000)


# ---------------------------------------------------------------- tuning never changes results

def test_every_ntt_chunk_gives_the_same_switch():
    ctx, ref = make(C2_SHAPE_N4096)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=31)
    polys = [ctx.uniform(l, seed=40 + l) for l in (L, L // 2, 3, 0)]
    want = None
    for chunk in (0, 1, 3, 8, 64):
        ctx.set_tuning("ntt_chunk_fwd", chunk)
        ctx.set_tuning("ntt_chunk_inv", chunk)
        got = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
        want = want or got
        assert got == want, chunk
    rep = ctx.autotune()
    assert rep["ntt_chunk_fwd"] in (0, 16, 24, 32, 48, 64, 96, 128)
    got = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
    assert got == want
    with pytest.raises(ckh.InvalidParams):
        ctx.set_tuning("no_such_knob", 1)


def test_graph_replays_equal_direct_launches():
    """Repeated batched switches of one level composition are captured once and replayed with new
    descriptors; every replay (new inputs, eager work and synchronizations in between) must equal
    the directly launched result."""
    ctx, ref = make(C2_SHAPE_N4096)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=51)
    lv = [L, 3, L // 2, 0, 7]
    ctx.set_tuning("graphs", 0)
    rounds = []
    for r in range(4):
        polys = [ctx.uniform(l, seed=500 + 10 * r + i) for i, l in enumerate(lv)]
        rounds.append((polys, [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]))
    c0, l0 = ctx.graph_stats()
    assert l0 == 0
    ctx.set_tuning("graphs", 1)
    for r, (polys, want) in enumerate(rounds):
        got = ctx.key_switch_batch(polys, key)
        # eager work on the same stream between replays
        extra = ctx.key_switch(polys[1], key)
        ctx.synchronize()
        assert [(kb.digest(), ka.digest()) for kb, ka in got] == want, r
        assert extra[0].digest() == want[1][0]
    caps, launches = ctx.graph_stats()
    assert caps >= 1 and launches >= 3
    # a replay word for word against the oracle
    polys, _ = rounds[3]
    kb, ka = ctx.key_switch_batch(polys, key)[2]
    d = polys[2].download()
    wb, wa = ref.key_switch(d.ravel(), lv[2], ref.uniform_key(L, 51), L)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()


@pytest.mark.parametrize("case", [C2_SHAPE_N4096, SMALL_N8192, C4])
def test_schedules_give_the_same_switch(case):
    """Every conversion kernel variant with and without the row pass fused into the inner product
    (and with it, the mod_down finish fused or not; mod_down's conversions on each variant of
    their own; the inverse NTTs in the conversions' y form or not) gives the same batched switch,
    equal to the oracle."""
    ctx, ref = make(case)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=61)
    lv = sorted({L, L // 2, 1, 0, max(L - 3, 0)})
    polys = [ctx.uniform(l, seed=600 + i) for i, l in enumerate(lv)]
    want = None
    for var, tg in ((0, 8), (1, 8), (0, 3), (1, 24), (2, 8), (0, 24)):
        for fuse in (0, 1):
            ctx.set_tuning("conv_tg", tg)
            ctx.set_tuning("conv_variant", var)
            ctx.set_tuning("fuse_ip", fuse)
            got = [(kb.download(), ka.download()) for kb, ka in ctx.key_switch_batch(polys, key)]
            if want is None:
                want = got
            for (wb, wa), (gb, ga) in zip(want, got):
                assert (wb == gb).all() and (wa == ga).all(), (var, fuse)
    ctx.set_tuning("conv_tg", 8)
    ctx.set_tuning("conv_variant", 0)
    ctx.set_tuning("fuse_ip", 1)
    for fin in (0, 1):
        for md in (0, 1, 2, 3):
            ctx.set_tuning("fuse_fin", fin)
            ctx.set_tuning("conv_variant_md", md)
            assert ctx.tuning("fuse_fin") == fin and ctx.tuning("conv_variant_md") == md
            got = [(kb.download(), ka.download()) for kb, ka in ctx.key_switch_batch(polys, key)]
            for (wb, wa), (gb, ga) in zip(want, got):
                assert (wb == gb).all() and (wa == ga).all(), ("fin", fin, "md", md)
    ctx.set_tuning("fuse_fin", 1)
    ctx.set_tuning("conv_variant_md", 1)
    # the inverse NTT in the conversions' y form on every subset of the conversion kinds, on the
    # vector and the matrix-core kernels
    for var in (0, 2):
        ctx.set_tuning("conv_variant", var)
        for yf in range(8):
            ctx.set_tuning("conv_yform", yf)
            assert ctx.tuning("conv_yform") == yf
            got = [(kb.download(), ka.download()) for kb, ka in ctx.key_switch_batch(polys, key)]
            for (wb, wa), (gb, ga) in zip(want, got):
                assert (wb == gb).all() and (wa == ga).all(), ("var", var, "conv_yform", yf)
    ctx.set_tuning("conv_variant", 0)
    ctx.set_tuning("conv_yform", 7)
    ob, oa = ref.key_switch(polys[-1].download().ravel(), lv[-1], ref.uniform_key(L, 61), L)
    assert (want[-1][0] == ob).all() and (want[-1][1] == oa).all()
    # the small-batch schedule: the whole batch (and each single switch) under every choice
    ctx.set_tuning("small_limbs", 1 << 20)
    for var, tg, fuse in ((0, 8, 0), (1, 4, 1), (2, 8, 0), (0, 24, 1)):
        ctx.set_tuning("conv_variant_small", var)
        ctx.set_tuning("conv_tg_small", tg)
        ctx.set_tuning("fuse_ip_small", fuse)
        got = [(kb.download(), ka.download()) for kb, ka in ctx.key_switch_batch(polys, key)]
        for (wb, wa), (gb, ga) in zip(want, got):
            assert (wb == gb).all() and (wa == ga).all(), ("small", var, tg, fuse)
        kb, ka = ctx.key_switch(polys[0], key)
        assert (kb.download() == want[0][0]).all() and (ka.download() == want[0][1]).all()


@pytest.mark.parametrize("case", [SMALL, C2_SHAPE_N4096, SMALL_N8192])
def test_matrix_core_conversion_equals_oracle(case):
    """The int8 matrix-core base conversion (variant 2) on every stage that converts: decompose,
    mod_down with near ties, rescale, and the full switch, word for word with the oracle."""
    ctx, ref = make(case)
    ctx.set_tuning("conv_variant", 2)
    L = len(ref.q) - 1
    for level in levels_for(len(ref.q)):
        a = host_uniform(ref, level, 71 + level)
        digs = ctx.decompose(ctx.upload(a, level))
        want = ref.decompose(a.ravel(), level)
        for j, d in enumerate(digs):
            assert (d.download() == want[j]).all(), (level, j)
        for ppart in (O.uniform_poly(ref.p, ref.n, 81 + level), near_tie_p_part(ref)):
            qp = np.concatenate([a, ppart])
            got = ctx.upload(qp, level, p_limbs=ref.k).mod_down().download()
            assert (got == ref.mod_down(a.ravel(), level, ppart.ravel())).all(), level
        if level >= 1:
            got = ctx.upload(a, level).rescale_by_last().download()
            assert (got == ref.rescale_by_last(a.ravel(), level)).all(), level
    key = ctx.ksk_uniform(L, seed=91)
    d = host_uniform(ref, L, 92)
    kb, ka = ctx.key_switch(ctx.upload(d, L), key)
    wb, wa = ref.key_switch(d.ravel(), L, ref.uniform_key(L, 91), L)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()


@pytest.mark.skip(reason="the stream-split batch path is switched off: its autotuning run (back-to-back "
                         "split batches on 4 streams at N = 2^16) faulted the GPU with a memory aperture "
                         "violation whose cause is not shown; the code path was removed")
def test_batch_split_over_streams_equals_one_stream():
    """A batched switch split over 2 and 3 streams (concurrent parts joined back into the caller's
    stream) equals the unsplit batch, on the first run and on graph replays."""
    ctx, ref = make(C2_SHAPE_N4096, n_streams=3)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=71)
    lv = [L, 3, L // 2, 0, 7, L - 1, 5, 1, L, 2]
    polys = [ctx.uniform(l, seed=700 + i) for i, l in enumerate(lv)]
    want = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
    for split in (2, 3):
        ctx.set_tuning("ks_split", split)
        for rep in range(3):
            got = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
            assert got == want, (split, rep)
    wb, wa = ref.key_switch(polys[2].download().ravel(), lv[2], ref.uniform_key(L, 71), L)
    kb, ka = ctx.key_switch_batch(polys, key)[2]
    assert (kb.download() == wb).all() and (ka.download() == wa).all()


def test_graph_replays_back_to_back_without_host_sync():
    """Replays of one captured plan enqueued back to back (new inputs each time, no host
    synchronization between them, every output kept) each equal the directly launched switch:
    a replay's descriptors must not reach the device before the previous replay is done."""
    ctx, ref = make(C2_SHAPE_N4096)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=81)
    lv = [L, 2, L // 2, 0, 9]
    rounds = [[ctx.uniform(l, seed=800 + 10 * r + i) for i, l in enumerate(lv)] for r in range(6)]
    ctx.set_tuning("graphs", 0)
    want = [[(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(p, key)] for p in rounds]
    ctx.set_tuning("graphs", 1)
    outs = [ctx.key_switch_batch(p, key) for p in rounds]  # first sighting, capture, 4 replays
    ctx.synchronize()
    got = [[(kb.digest(), ka.digest()) for kb, ka in o] for o in outs]
    assert got == want
    assert ctx.graph_stats()[1] >= 4



P_F64_N4096 = ("p45_n4096", 12, [40] * 4, [45] * 2, 2)  # special primes on the f64 path


@pytest.mark.parametrize("case", [C2_SHAPE_N4096, SMALL_N8192, P_F64_N4096], ids=lambda c: c[0])
def test_inverse_chunks_and_special_prime_sizes_give_the_same_switch(case):
    """Batched switches and hoisted rotations (digits read through sigma_g) with the mod_down
    INTT in chunks of 3 limbs or in one launch, with 60-bit and with 45-bit special primes (the
    latter on the f64 path of the register-radix NTT), equal to each other and to the oracle."""
    say("inverse chunks", case[0])
    ctx, ref = make(case)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=71)
    lv = sorted({L, L // 2, 1, 0})
    polys = [ctx.uniform(l, seed=700 + i) for i, l in enumerate(lv)]
    steps = [1, 5]
    gkeys = [ctx.ksk_uniform(L, seed=80 + s, galois=ckh.galois_elt(ref.log_n, s)) for s in steps]
    ev = ckh.Evaluator(ctx, None, gkeys)
    ct = ctx.ct_from_polys([ctx.uniform(L - 1, seed=750), ctx.uniform(L - 1, seed=751)], 2.0 ** 30)
    res = []
    for chunk in (3, 0):
        ctx.set_tuning("ntt_chunk_inv", chunk)
        ks = [(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(polys, key)]
        hr = [tuple(r.digests()) for r in ev.hoisted_rotate(ct, steps)]
        res.append((ks, hr))
    assert res[0] == res[1]
    kb, ka = ctx.key_switch_batch(polys, key)[-1]
    wb, wa = ref.key_switch(polys[-1].download().ravel(), lv[-1], ref.uniform_key(L, 71), L)
    assert (kb.download() == wb).all() and (ka.download() == wa).all()
    c0, c1 = ct.download()
    o = ref.apply_galois(c0.ravel(), c1.ravel(), L - 1, ckh.galois_elt(ref.log_n, 5), ref.uniform_key(L, 85), L)
    got = ev.hoisted_rotate(ct, [5])[0].download()
    assert (got[0] == o[0]).all() and (got[1] == o[1]).all()


def test_plan_eviction_under_churn_keeps_results():
    """A plan cache of two entries under five alternating batch shapes, with no host wait between
    calls: plans are evicted while their last replay may still run (destroyed once it finishes),
    and every result equals the directly launched one."""
    say("plan eviction churn")
    ctx, ref = make(C2_SHAPE_N4096)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=91)
    shapes = [[L], [L, 3], [5, 5, 5], [L // 2, 1], [0, L, 2]]
    batches = [[ctx.uniform(l, seed=900 + 10 * i + j) for j, l in enumerate(sh)] for i, sh in enumerate(shapes)]
    ctx.set_tuning("graphs", 0)
    want = [[(kb.digest(), ka.digest()) for kb, ka in ctx.key_switch_batch(b, key)] for b in batches]
    ctx.set_tuning("graphs", 1)
    ctx.set_tuning("graph_cache", 2)
    outs = []
    for r in range(3):
        for i, b in enumerate(batches):
            for _ in range(3):  # seen, captured, replayed; then evicted by the next shapes
                outs.append((i, ctx.key_switch_batch(b, key)))
    ctx.synchronize()
    for i, o in outs:
        assert [(kb.digest(), ka.digest()) for kb, ka in o] == want[i], i
    caps, launches = ctx.graph_stats()
    assert caps >= 15 and launches >= 15


@pytest.mark.parametrize("dnum", [2, 1], ids=["alpha8", "alpha16"])
def test_wide_digits_key_switch(dnum):
    """Digits of 8 and 16 primes (the conversion's 8- and 16-source kernels, whose 60-bit special
    targets take the exact 128-bit dot product) against the oracle, batched and single."""
    say("wide digits", dnum)
    case = ("wide", 10, [40] * 16, [60] * 3, dnum)
    ctx, ref = make(case)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=95)
    hkey = ref.uniform_key(L, 95)
    lv = [L, 9, 3]
    polys = [ctx.uniform(l, seed=950 + l) for l in lv]
    got = ctx.key_switch_batch(polys, key)
    for p_, l, (kb, ka) in zip(polys, lv, got):
        wb, wa = ref.key_switch(p_.download().ravel(), l, hkey, L)
        assert (kb.download() == wb).all() and (ka.download() == wa).all(), l
    kb, ka = ctx.key_switch(polys[0], key)
    assert kb.digest() == got[0][0].digest() and ka.digest() == got[0][1].digest()
