"""Shared parameter sets and seeded inputs for the device parity tests: the device library and
the CPU oracle get the same primes and the same seeded residues, and results are compared word
for word."""
import numpy as np

import ckh
import oracle as O

# (name, log_n, q_bits, p_bits, dnum)
SMALL = ("small_a2", 10, [60, 40, 40, 40, 40, 40], [60, 60], 3)     # alpha 2, two special primes
ALPHA1 = ("alpha1", 10, [60, 40, 40, 40], [60], 4)                  # one-prime digits, k = 1
C2_SHAPE_SMALL_N = ("c2_shape_n1024", 10, [50] * 24, [60] * 6, 4)   # C2 prime shape at N = 2^10
C4_SHAPE_SMALL_N = ("c4_shape_n1024", 10, [50] * 20, [60] * 4, 5)   # C4 prime shape at N = 2^10
C2_SHAPE_N4096 = ("c2_shape_n4096", 12, [50] * 24, [60] * 6, 4)   # register-radix NTT path
SMALL_N8192 = ("small_a3_n8192", 13, [50] * 6, [60] * 2, 2)        # alpha 3, odd log N split
C2 = ("c2", 16, [50] * 24, [60] * 6, 4)                             # the headline key switch
C4 = ("c4", 15, [50] * 20, [60] * 4, 5)


def make(case, n_streams=1):
    name, log_n, qb, pb, dnum = case
    q, p = ckh.params_primes(log_n, qb, pb)
    ctx = ckh.Context(log_n, q, p, dnum, n_streams=n_streams)
    ref = O.Params(log_n, q=q, p=p, dnum=dnum)
    assert ref.q == q and ref.p == p
    return ctx, ref


def levels_for(nq):
    """Levels spread over top / top-1 / 3L/4 / L/2 / L/3 / 2 / 1 / 0 (clipped digits included)."""
    L = nq - 1
    return sorted({L, max(L - 1, 0), 3 * L // 4, L // 2, L // 3, min(2, L), min(1, L), 0}, reverse=True)


def host_uniform(ref, level, seed, stream0=0, p_limbs=0):
    primes = ref.q[: level + 1] + (ref.p if p_limbs else [])
    return O.uniform_poly(primes, ref.n, seed, stream0)


def near_tie_digits(ref, level):
    """Coefficient-form limbs where every digit D puts [x + h]_D next to 0 or Q_D (the f64
    estimate of the exact conversion is then decided by its last bits), transformed to NTT form
    by the oracle."""
    n = ref.n
    limbs = np.zeros((level + 1, n), dtype=np.uint64)
    for s, e in ref.digit_ranges(level):
        D = 1
        for q in ref.q[s:e]:
            D *= q
        for i in range(s, e):
            r = ref.q[i]
            for c in range(n):
                m = c // 2 + 1
                sv = m % r if c % 2 == 0 else (r - m % r) % r
                limbs[i, c] = (sv - (r >> 1)) % r
    return O.ntt_forward(n, ref.q[: level + 1], limbs.ravel()).reshape(level + 1, n)


def near_tie_p_part(ref):
    """P limbs (coefficient form) with [X + h]_P next to 0 or P in every coefficient."""
    n = ref.n
    P = 1
    for p in ref.p:
        P *= p
    limbs = np.zeros((len(ref.p), n), dtype=np.uint64)
    for t, r in enumerate(ref.p):
        for c in range(n):
            m = c // 2 + 1
            sv = m % r if c % 2 == 0 else (r - m % r) % r
            limbs[t, c] = (sv - (r >> 1)) % r
    return O.ntt_forward(n, ref.p, limbs.ravel()).reshape(len(ref.p), n)


def oracle_bsgs(ref, c0, c1, level, present, baby, giant, stride, offset, mul, finish, share_giant,
                key_of_step, kml):
    """The BSGS driver composed from oracle operations, in the reference's order
    (crates/ckks-train-ckks/src/linalg.rs:800-888): hoisted baby rotations of the used steps;
    per giant block k1 with shift (offset + baby k1) stride: with share_giant (EvalConfig::bsgs),
    finish(sum over k2 of mul(i, rot(x, k2 s))) rotated by the shift, else (shift != 0) every
    product finished and rotated by itself; blocks summed in k1 order. mul(i, (x0, x1)) returns
    the components of entry i times the rotated x; key_of_step(step) the host key of a step."""
    slots = ref.n // 2
    used = [k2 for k2 in range(baby) if any(present[k1 * baby + k2] for k1 in range(giant))]
    nontriv = [k2 for k2 in used if (k2 * stride) % slots]
    elts = [galois_elt_host(ref.n, k2 * stride) for k2 in nontriv]
    hoisted = ref.hoisted_galois(c0, c1, level, elts, [key_of_step(k2 * stride) for k2 in nontriv], kml)
    xs = {}
    for k2 in used:
        if k2 in nontriv:
            i = nontriv.index(k2)
            xs[k2] = (hoisted[i, 0], hoisted[i, 1])
        else:
            xs[k2] = (c0.reshape(level + 1, ref.n), c1.reshape(level + 1, ref.n))

    def add(a, b):
        return b if a is None else tuple(ref.poly_add(u, v, level) for u, v in zip(a, b))

    def rot(t, shift):
        if shift % slots == 0:
            return t
        return ref.apply_galois(t[0], t[1], level, galois_elt_host(ref.n, shift), key_of_step(shift % slots), kml)

    acc = None
    for k1 in range(giant):
        shift = (offset + baby * k1) * stride
        if not share_giant and shift % slots:
            for k2 in range(baby):
                if present[k1 * baby + k2]:
                    acc = add(acc, rot(finish(mul(k1 * baby + k2, xs[k2])), shift))
            continue
        blk = None
        for k2 in range(baby):
            if present[k1 * baby + k2]:
                blk = add(blk, mul(k1 * baby + k2, xs[k2]))
        if blk is not None:
            acc = add(acc, rot(finish(blk), shift))
    return acc


def oracle_matvec_sum(ref, xs, level, entries, baby, giant, stride, offset, share_giant, key_of_step, kml):
    """matvec_plain_sum composed from oracle operations (linalg.rs:800-888 with several terms):
    every term's baby rotations hoisted; per block the products of every term summed (or, naive
    and a nonzero shift, each rotated by itself); xs[t] = (c0, c1), entries[t][i] a host array
    or None."""
    slots = ref.n // 2
    rx = []
    for (c0, c1), ents in zip(xs, entries):
        used = [k2 for k2 in range(baby) if any(ents[k1 * baby + k2] is not None for k1 in range(giant))]
        nontriv = [k2 for k2 in used if (k2 * stride) % slots]
        hoisted = ref.hoisted_galois(c0, c1, level, [galois_elt_host(ref.n, k2 * stride) for k2 in nontriv],
                                     [key_of_step(k2 * stride) for k2 in nontriv], kml)
        row = {}
        for k2 in used:
            if k2 in nontriv:
                i = nontriv.index(k2)
                row[k2] = (hoisted[i, 0], hoisted[i, 1])
            else:
                row[k2] = (c0.reshape(level + 1, ref.n), c1.reshape(level + 1, ref.n))
        rx.append(row)

    def add(a, b):
        return b if a is None else tuple(ref.poly_add(u, v, level) for u, v in zip(a, b))

    def rot(t, shift):
        if shift % slots == 0:
            return t
        return ref.apply_galois(t[0], t[1], level, galois_elt_host(ref.n, shift), key_of_step(shift % slots), kml)

    def mul(t, i, x):
        return tuple(ref.poly_mul(xc, entries[t][i], level) for xc in x)

    acc = None
    for k1 in range(giant):
        shift = (offset + baby * k1) * stride
        blk = None
        for t in range(len(xs)):
            for k2 in range(baby):
                i = k1 * baby + k2
                if entries[t][i] is None:
                    continue
                if not share_giant and shift % slots:
                    acc = add(acc, rot(mul(t, i, rx[t][k2]), shift))
                else:
                    blk = add(blk, mul(t, i, rx[t][k2]))
        if blk is not None:
            acc = add(acc, rot(blk, shift))
    return acc


def oracle_matvec(ref, c0, c1, level, entries, baby, giant, stride, key_of_step, kml, offset=0, share_giant=True):
    """matvec_plain composed from oracle operations: mul = mul_plain by the host entry."""
    present = [e is not None for e in entries]
    return oracle_bsgs(ref, c0, c1, level, present, baby, giant, stride, offset,
                       lambda i, x: tuple(ref.poly_mul(xc, entries[i], level) for xc in x),
                       lambda t: t, share_giant, key_of_step, kml)


def galois_elt_host(n, step):
    return O.galois_elt(n, step)


def c4_entries_mask(baby=32, giant=32):
    """The C4 product: in giant block k1 the diagonals at baby steps k1 and k1 + 16 (mod 32)."""
    return [[(k2 == k1 % baby) or (k2 == (k1 + baby // 2) % baby) for k2 in range(baby)] for k1 in range(giant)]


def oracle_linear_stage(ref, xs, level, entries, shape, key_of_step, kml):
    """apply_stage (linear.rs:900-919) from oracle operations: the BSGS sum of the terms
    (xs[t] = (c0, c1) with entries[t]), then the rescale of both components."""
    baby, giant, stride, offset = shape
    y = oracle_matvec_sum(ref, xs, level, entries, baby, giant, stride, offset, True, key_of_step, kml)
    r = ref.rescale_ct(np.stack([np.asarray(c).reshape(level + 1, ref.n) for c in y]), level)
    return r[0], r[1]


def oracle_coeff_to_slot(ref, c0, c1, stages, key_of_step, conj_key, kml):
    """coeff_to_slot (linear.rs:776-806) from oracle operations; stages are (level, entries,
    second entries or None, shape)."""
    conj = 2 * ref.n - 1
    x = (c0, c1)
    for level, ents, sec, shape in stages:
        xs, es = [x], [ents]
        if sec is not None:
            xs.append(ref.apply_galois(x[0], x[1], level, conj, conj_key, kml))
            es.append(sec)
        x = oracle_linear_stage(ref, xs, level, es, shape, key_of_step, kml)
    level = stages[-1][0] - 1
    xc = ref.apply_galois(x[0], x[1], level, conj, conj_key, kml)
    re = tuple(ref.poly_add(a, b, level) for a, b in zip(x, xc))
    im = tuple(ref.mul_by_x_half(ref.poly_sub(b, a, level), level) for a, b in zip(x, xc))
    return re, im


def oracle_slot_to_coeff(ref, re, im, stages, key_of_step, kml):
    """slot_to_coeff (linear.rs:824-859) from oracle operations."""
    level, ents, sec, shape = stages[0]
    x = oracle_linear_stage(ref, [re, im], level, [ents, sec], shape, key_of_step, kml)
    for level, ents, _, shape in stages[1:]:
        x = oracle_linear_stage(ref, [x], level, [ents], shape, key_of_step, kml)
    return x


def bsgs_rotation_count(masks, baby, giant, stride, offset, slots):
    """Rotations of one BSGS sum (linalg.rs:695-790 docs): per term, a baby rotation for each
    nonzero baby step a present entry of that term uses; a giant rotation for each block with a
    present entry of any term and a nonzero shift."""
    n = 0
    for m in masks:
        n += sum(1 for k2 in range(1, baby)
                 if (k2 * stride) % slots and any(m[k1 * baby + k2] for k1 in range(giant)))
    n += sum(1 for k1 in range(giant)
             if ((offset + baby * k1) * stride) % slots and any(m[k1 * baby + k2] for m in masks for k2 in range(baby)))
    return n
