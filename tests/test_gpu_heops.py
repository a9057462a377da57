"""The HeOps surface beyond key switching against the oracle, bit for bit, through the C ABI:
mul_scalar / mul_plain_to_scale with their exact labels (evaluator.rs:330-418,738-797), the
EvalConfig flags (ops.rs:111-134: hoist, bsgs, lazy_relin), mul_lazy / relin_lazy
(ops.rs:339-356), and the BSGS products with a diagonal offset and encrypted diagonals
(linalg.rs:608-665,800-888)."""
import numpy as np
import pytest

import ckh
import oracle as O
from gpu_cases import C2, C4_SHAPE_SMALL_N, SMALL, galois_elt_host, host_uniform, make, oracle_bsgs, oracle_matvec

pytestmark = pytest.mark.gpu

DELTA = 2.0 ** 40


def _ct(ctx, ref, level, seed, ncomp=2, scale=DELTA):
    polys = np.stack([host_uniform(ref, level, seed + i) for i in range(ncomp)])
    return ctx.ct_upload(polys, level, scale), polys


@pytest.fixture(scope="module", params=[SMALL, C4_SHAPE_SMALL_N], ids=lambda c: c[0])
def env(request):
    print("case", request.param[0], flush=True)
    ctx, ref = make(request.param)
    yield ctx, ref
    ctx.synchronize()


def test_exact_product_scale_matches_oracle(env):
    ctx, ref = env
    ev = ckh.Evaluator(ctx)
    rng = np.random.default_rng(3)
    for level in range(1, len(ref.q)):
        for t in [DELTA, 1.5 * DELTA, DELTA * DELTA / ref.q[level], 3.0] + list(2.0 ** rng.uniform(20, 60, 20)):
            try:
                want = ref.exact_product_scale(level, float(t))
            except ValueError as e:
                with pytest.raises(getattr(ckh, str(e))):
                    ev.exact_product_scale(level, float(t))
                continue
            assert ev.exact_product_scale(level, float(t)) == want
    with pytest.raises(ckh.NoLevelsLeft):
        ev.exact_product_scale(0, DELTA)
    with pytest.raises(ckh.InvalidParams):
        ev.exact_product_scale(len(ref.q), DELTA)
    with pytest.raises(ckh.InvalidParams):
        ev.exact_product_scale(1, -1.0)


def test_mul_scalar_matches_oracle_and_lands_on_the_target(env):
    ctx, ref = env
    L = len(ref.q) - 1
    ev = ckh.Evaluator(ctx)
    for level in (L, 1):
        for ncomp in (2, 3):
            ct, h = _ct(ctx, ref, level, 700 + 10 * level + ncomp, ncomp)
            for target in (DELTA, 1.5 * DELTA):
                for value in (1.0, 0.5, -1.25, 1e-3, 0.0125, 3.0, 0.0):
                    ev.reset_counter()
                    got = ev.mul_scalar(ct, value, target)
                    label, _, k = ref.mul_scalar_parts(level, DELTA, value, target)
                    assert got.scale == label and got.level == level and got.size() == ncomp
                    d = got.download()
                    for c in range(ncomp):
                        want = ref.poly_mul_scalar(h[c].ravel(), level, int(abs(k)), k < 0)
                        assert (d[c] == want).all(), (level, ncomp, target, value)
                    assert ev.counter().ct_pt_mults == 1
                    if ncomp == 2:
                        assert ev.rescale(got).scale == target
    ct0, _ = _ct(ctx, ref, 0, 790)
    with pytest.raises(ckh.NoLevelsLeft):
        ev.mul_scalar(ct0, 1.0, DELTA)
    ct, _ = _ct(ctx, ref, L, 791)
    for bad in (float("nan"), float("inf")):
        with pytest.raises(ckh.InvalidParams):
            ev.mul_scalar(ct, bad, DELTA)
    for bad in (0.0, -DELTA, float("nan"), float("inf")):
        with pytest.raises(ckh.InvalidParams):
            ev.mul_scalar(ct, 1.0, bad)
    with pytest.raises(ckh.InvalidParams):  # an integer at or above 2^62
        ev.mul_scalar(ct, 1e9, DELTA * 1e9)


def test_mul_plain_to_scale(env):
    ctx, ref = env
    L = len(ref.q) - 1
    ev = ckh.Evaluator(ctx)
    ct, h = _ct(ctx, ref, L, 800)
    target = DELTA
    sp = ev.target_plain_scale(ct, target)
    assert sp == ref.exact_product_scale(L, target) / DELTA
    w = host_uniform(ref, L, 810)
    pt = ctx.upload(w, L)
    got = ev.mul_plain_to_scale(ct, pt, sp, target)
    assert got.scale == ref.exact_product_scale(L, target)
    d = got.download()
    for c in range(2):
        assert (d[c] == ref.poly_mul(h[c], w, L)).all()
    assert ev.rescale(got).scale == target
    with pytest.raises(ckh.ScaleMismatch):
        ev.mul_plain_to_scale(ct, pt, sp * 2.0, target)
    with pytest.raises(ckh.LevelMismatch):
        ev.mul_plain_to_scale(ct, ctx.upload(host_uniform(ref, L - 1, 811), L - 1), sp, target)
    with pytest.raises(ckh.InvalidParams):  # plaintext scale below 2^20
        ev.target_plain_scale(ct, 1.0)


def test_config_flags_and_lazy_products(env):
    ctx, ref = env
    L = len(ref.q) - 1
    relin = ctx.ksk_uniform(L, seed=14)
    steps = [1, 3, -2]
    gkeys = [ctx.ksk_uniform(L, seed=60 + i, galois=ckh.galois_elt(ref.log_n, s)) for i, s in enumerate(steps)]
    ev = ckh.Evaluator(ctx, relin, gkeys)
    assert ev.config() == {"hoist": True, "bsgs": True, "lazy_relin": True}
    a, ha = _ct(ctx, ref, L, 900)
    b, hb = _ct(ctx, ref, L, 910)
    # hoist off: individual rotations, same words and counts
    hoisted = [o.download() for o in ev.hoisted_rotate(a, steps)]
    c_h = ev.counter().rotations
    ev.set_config(hoist=False)
    assert ev.config()["hoist"] is False
    ev.reset_counter()
    single = [o.download() for o in ev.hoisted_rotate(a, steps)]
    assert ev.counter().rotations == c_h == 3
    assert all((x == y).all() for x, y in zip(hoisted, single))
    # lazy_relin on: mul_lazy is the tensor product, relin_lazy relinearizes
    ev.set_config()
    t = ev.mul_lazy(a, b)
    assert t.size() == 3
    d0, d1, d2 = ref.mul_no_relin(ha[0], ha[1], hb[0], hb[1], L)
    got = t.download()
    assert (got[0] == d0).all() and (got[1] == d1).all() and (got[2] == d2).all()
    r = ev.relin_lazy(t).download()
    w0, w1 = ref.relinearize(d0, d1, d2, L, ref.uniform_key(L, 14), L)
    assert (r[0] == w0).all() and (r[1] == w1).all()
    # lazy_relin off: mul_lazy is mul, relin_lazy passes its input through
    ev.set_config(lazy_relin=False)
    m = ev.mul_lazy(a, b)
    assert m.size() == 2
    got = m.download()
    assert (got[0] == w0).all() and (got[1] == w1).all()
    assert (ev.relin_lazy(m).download() == got).all()
    # fork keeps the configuration
    assert ev.fork().config() == {"hoist": True, "bsgs": True, "lazy_relin": False}


def _keys_for(ctx, ref, L, steps, seed):
    slots = ref.n // 2
    norm = sorted({s % slots for s in steps if s % slots})
    keys, hseed = [], {}
    for i, st in enumerate(norm):
        keys.append(ctx.ksk_uniform(L, seed=seed + 10 * i, galois=ckh.galois_elt(ref.log_n, st)))
        hseed[st] = seed + 10 * i
    cache = {}

    def key_of_step(st):
        st %= slots
        if st not in cache:
            cache[st] = ref.uniform_key(L, hseed[st])
        return cache[st]

    return keys, key_of_step


@pytest.mark.parametrize("offset,bsgs", [(0, True), (-3, True), (-3, False), (0, False)])
def test_matvec_plain_offset_and_naive(offset, bsgs):
    ctx, ref = make(C4_SHAPE_SMALL_N)
    L = len(ref.q) - 1
    baby, giant, stride = 4, 4, 2
    steps = [k2 * stride for k2 in range(1, baby)] + [(offset + baby * k1) * stride for k1 in range(giant)]
    keys, key_of_step = _keys_for(ctx, ref, L, steps, 1200)
    ev = ckh.Evaluator(ctx, None, keys)
    ev.set_config(bsgs=bsgs)
    level = L - 1
    c0, c1 = host_uniform(ref, level, 1201), host_uniform(ref, level, 1202)
    x = ctx.ct_upload(np.stack([c0, c1]), level, DELTA)
    mask = [(i % 5) != 2 for i in range(baby * giant)]
    entries_h = [host_uniform(ref, level, 1300 + i) if m else None for i, m in enumerate(mask)]
    entries_d = [ctx.upload(w, level) if w is not None else None for w in entries_h]
    got = ev.matvec_plain(x, entries_d, DELTA, baby, giant, stride, offset).download()
    want = oracle_matvec(ref, c0, c1, level, entries_h, baby, giant, stride, key_of_step, L, offset=offset,
                         share_giant=bsgs)
    assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
    with pytest.raises(ckh.InvalidParams):
        ev.matvec_plain(x, entries_d, DELTA, baby, giant, stride, 1)
    with pytest.raises(ckh.InvalidParams):
        ev.matvec_plain(x, entries_d, DELTA, baby, giant, stride, -baby * giant)


@pytest.mark.parametrize("lazy,bsgs", [(True, True), (False, True), (True, False)])
def test_matvec_enc_matches_oracle(lazy, bsgs):
    ctx, ref = make(C4_SHAPE_SMALL_N)
    L = len(ref.q) - 1
    baby, giant, stride, offset = 4, 2, 1, -1
    steps = [k2 * stride for k2 in range(1, baby)] + [(offset + baby * k1) * stride for k1 in range(giant)]
    keys, key_of_step = _keys_for(ctx, ref, L, steps, 1400)
    relin = ctx.ksk_uniform(L, seed=1450)
    hrelin = ref.uniform_key(L, 1450)
    ev = ckh.Evaluator(ctx, relin, keys)
    ev.set_config(bsgs=bsgs, lazy_relin=lazy)
    level = L - 2
    c0, c1 = host_uniform(ref, level, 1401), host_uniform(ref, level, 1402)
    x = ctx.ct_upload(np.stack([c0, c1]), level, DELTA)
    ents_h, ents_d = [], []
    for i in range(baby * giant):
        if i == 5:
            ents_h.append(None)
            ents_d.append(None)
            continue
        e = np.stack([host_uniform(ref, level, 1500 + 2 * i), host_uniform(ref, level, 1501 + 2 * i)])
        ents_h.append(e)
        ents_d.append(ctx.ct_upload(e, level, DELTA))
    ev.reset_counter()
    got = ev.matvec_enc(x, ents_d, baby, giant, stride, offset)
    assert got.size() == 2 and got.level == level and got.scale == DELTA * DELTA

    def mul(i, xr):
        d = ref.mul_no_relin(ents_h[i][0], ents_h[i][1], xr[0], xr[1], level)
        if lazy:
            return tuple(d)
        return ref.relinearize(d[0], d[1], d[2], level, hrelin, L)

    def finish(t):
        return ref.relinearize(t[0], t[1], t[2], level, hrelin, L) if lazy else t

    present = [e is not None for e in ents_h]
    want = oracle_bsgs(ref, c0, c1, level, present, baby, giant, stride, offset, mul, finish, bsgs, key_of_step, L)
    g = got.download()
    assert (g[0] == want[0]).all() and (g[1] == want[1]).all()
    c = ev.counter()
    n_products = sum(present)
    assert c.ct_ct_tensor == n_products
    if lazy and bsgs:
        assert c.relins == giant  # one per block
    elif not lazy:
        assert c.relins == n_products
    x0 = ctx.ct_upload(np.stack([host_uniform(ref, 0, 1), host_uniform(ref, 0, 2)]), 0, DELTA)
    ents0 = [ctx.ct_upload(np.stack([host_uniform(ref, 0, 3), host_uniform(ref, 0, 4)]), 0, DELTA)] * (baby * giant)
    with pytest.raises(ckh.NoLevelsLeft):
        ev.matvec_enc(x0, ents0, baby, giant, stride, offset)
    with pytest.raises(ckh.InvalidParams):  # no relinearization key
        ckh.Evaluator(ctx, None, keys).matvec_enc(x, ents_d, baby, giant, stride, offset)
    bad = list(ents_d)
    bad[0] = ctx.ct_upload(ents_h[0], level, 2 * DELTA)
    with pytest.raises(ckh.ScaleMismatch):
        ev.matvec_enc(x, bad, baby, giant, stride, offset)


def _enc_job(ctx, ref, level, baby, giant, seed, missing=()):
    c0, c1 = host_uniform(ref, level, seed), host_uniform(ref, level, seed + 1)
    x = ctx.ct_upload(np.stack([c0, c1]), level, DELTA)
    ents = []
    for i in range(baby * giant):
        if i in missing:
            ents.append(None)
            continue
        e = np.stack([host_uniform(ref, level, seed + 10 + 2 * i), host_uniform(ref, level, seed + 11 + 2 * i)])
        ents.append(ctx.ct_upload(e, level, DELTA))
    return x, ents


def test_matvec_enc_batch_equals_individual_and_counts_only_on_success():
    """CudaEvaluator::matvec_enc_batch (crates/ckks-train-cuda/src/evaluator.rs:412-519) and the
    success-only counts of composite operations (src/matvec.rs `counted`): an allocation failure
    part-way, forced from the host, is a Device error that leaves the counter unchanged."""
    ctx, ref = make(C4_SHAPE_SMALL_N)
    L = len(ref.q) - 1
    shapes = [(4, 2, 1, 0), (2, 2, 2, -1), (4, 2, 1, -3)]
    steps = set()
    for baby, giant, stride, offset in shapes:
        steps |= {k2 * stride for k2 in range(1, baby)} | {(offset + baby * k1) * stride for k1 in range(giant)}
    keys, _ = _keys_for(ctx, ref, L, sorted(steps), 1600)
    ev = ckh.Evaluator(ctx, ctx.ksk_uniform(L, seed=1650), keys)
    jobs = [_enc_job(ctx, ref, lv, sh[0], sh[1], 1700 + 100 * j, missing=(1,) if j == 1 else ())
            for j, (lv, sh) in enumerate(zip((L - 1, 3, L), shapes))]
    single, counts = [], []
    for (x, ents), sh in zip(jobs, shapes):
        ev.reset_counter()
        single.append(ev.matvec_enc(x, ents, *sh).download())
        counts.append(ev.counter())
    ev.reset_counter()
    batch = ev.matvec_enc_batch([j[0] for j in jobs], [j[1] for j in jobs], shapes)
    for b, s_ in zip(batch, single):
        assert (b.download() == s_).all()
    c = ev.counter()
    for f in ("rotations", "relins", "ct_ct_tensor"):
        assert getattr(c, f) == sum(getattr(k, f) for k in counts), f
    assert ev.matvec_enc_workspace_bytes(L, 4, 2) > 8 * ref.n * (L + 1) * 2
    # an allocation failure part-way: Device error, nothing counted
    for hook_at in (3, 40):
        ev.reset_counter()
        ctx.set_tuning("fail_alloc_after", hook_at)
        with pytest.raises(ckh.Device):
            ev.matvec_enc_batch([j[0] for j in jobs], [j[1] for j in jobs], shapes)
        ctx.set_tuning("fail_alloc_after", -1)
        c = ev.counter()
        assert (c.rotations, c.relins, c.ct_ct_tensor) == (0, 0, 0), hook_at
    ev.reset_counter()
    ctx.set_tuning("fail_alloc_after", 25)
    with pytest.raises(ckh.Device):
        ev.matvec_enc(jobs[0][0], jobs[0][1], *shapes[0])
    ctx.set_tuning("fail_alloc_after", -1)
    assert ev.counter().rotations == 0 and ev.counter().relins == 0
    # and everything works again afterwards
    assert (ev.matvec_enc(jobs[0][0], jobs[0][1], *shapes[0]).download() == single[0]).all()
    with pytest.raises(ckh.InvalidParams):
        ev.matvec_enc_batch([jobs[0][0]], [], [shapes[0]])


def test_pool_accounting_and_trim():
    """CudaDevice::pool_used_memory / set_pool_release_threshold / trim_pool (device.rs:478-527)."""
    ctx, ref = make(SMALL)
    L = len(ref.q) - 1
    base = ctx.pool_used_memory()
    polys = [ctx.uniform(L, seed=i) for i in range(8)]
    ctx.synchronize()
    used = ctx.pool_used_memory()
    assert used >= base + 8 * (L + 1) * ref.n * 8
    del polys
    ctx.synchronize()
    assert ctx.pool_used_memory() < used
    ctx.trim_pool(0)  # the cached blocks back to the driver
    again = [ctx.uniform(L, seed=i) for i in range(8)]  # allocations after a trim work
    assert all(a.download().shape == (L + 1, ref.n) for a in again)
    del again
    # under a release threshold of zero freed blocks go straight back to the driver
    ctx.set_pool_release_threshold(0)
    q = ctx.uniform(L, seed=98)
    del q
    ctx.synchronize()
    ctx.set_pool_release_threshold((1 << 64) - 1)
    p = ctx.uniform(L, seed=99)
    assert p.download().shape == (L + 1, ref.n)


# ---------------------------------------------------------------- mod-raise (bootstrapping)

@pytest.mark.parametrize("case", [SMALL, C4_SHAPE_SMALL_N, C2], ids=lambda c: c[0])
def test_raise_from_level0_matches_oracle(case):
    ctx, ref = make(case)
    L = len(ref.q) - 1
    a = host_uniform(ref, 0, 2100)
    a[0, :4] = [0, 1, ref.q[0] // 2, ref.q[0] // 2 + 1]
    for ntt in (True, False):
        p = ctx.upload(a, 0, ntt=ntt)
        for top in sorted({0, 1, L // 2, L}):
            got = p.raise_from_level0(top).download()
            want = ref.raise_from_level0(a.ravel(), top, ntt=ntt)
            assert (got == want).all(), (ntt, top)
    with pytest.raises(ckh.InvalidParams):
        ctx.upload(host_uniform(ref, 1, 2101), 1).raise_from_level0(L)
    with pytest.raises(ckh.InvalidParams):
        ctx.upload(a, 0).raise_from_level0(L + 1)


@pytest.mark.parametrize("case", [SMALL, C4_SHAPE_SMALL_N, C2], ids=lambda c: c[0])
def test_mod_raise_with_sse_matches_oracle(case):
    """Evaluator::mod_raise (bootstrap/modraise.rs:53-69): switch to the ephemeral key at level 0,
    raise both components, switch back at the top; and mod_raise_main_key."""
    ctx, ref = make(case)
    L = len(ref.q) - 1
    to_sparse, to_dense = ctx.ksk_uniform(0, seed=2200), ctx.ksk_uniform(L, seed=2201)
    h_sparse, h_dense = ref.uniform_key(0, 2200), ref.uniform_key(L, 2201)
    ev = ckh.Evaluator(ctx)
    ct, h = _ct(ctx, ref, 0, 2210)
    got = ev.mod_raise(ct, to_sparse, to_dense)
    assert got.level == L and got.scale == DELTA and got.size() == 2
    kb, ka = ref.key_switch(h[1].ravel(), 0, h_sparse, 0)
    c0 = ref.poly_add(h[0], kb, 0)
    r0, r1 = ref.raise_from_level0(c0.ravel(), L), ref.raise_from_level0(ka.ravel(), L)
    kb2, ka2 = ref.key_switch(r1.ravel(), L, h_dense, L)
    g = got.download()
    assert (g[0] == ref.poly_add(r0, kb2, L)).all() and (g[1] == ka2).all()
    c = ev.counter()
    assert c.sse_switches == 2 and c.key_switches() == 2
    m = ev.mod_raise_main_key(ct).download()
    assert (m[0] == ref.raise_from_level0(h[0].ravel(), L)).all()
    assert (m[1] == ref.raise_from_level0(h[1].ravel(), L)).all()
    assert ev.counter().sse_switches == 2
    ct1, _ = _ct(ctx, ref, 1, 2220)
    with pytest.raises(ckh.LevelMismatch):
        ev.mod_raise(ct1, to_sparse, to_dense)
    ct3, _ = _ct(ctx, ref, 0, 2230, ncomp=3)
    with pytest.raises(ckh.InvalidParams):
        ev.mod_raise_main_key(ct3)
    with pytest.raises(ckh.InvalidParams):  # the keys swapped
        ev.mod_raise(ct, to_dense, to_sparse)
    assert ev.counter().sse_switches == 2


@pytest.mark.parametrize("bsgs", [True, False])
def test_matvec_plain_sum_matches_oracle(bsgs):
    """Evaluator::matvec_plain_sum_sparse (linalg.rs:729-790): two terms sharing the giant
    rotations, zero entries skipped (a column unused by one term costs that term no baby
    rotation)."""
    from gpu_cases import oracle_matvec_sum
    ctx, ref = make(C4_SHAPE_SMALL_N)
    L = len(ref.q) - 1
    baby, giant, stride, offset = 4, 4, 1, -2
    steps = [k2 * stride for k2 in range(1, baby)] + [(offset + baby * k1) * stride for k1 in range(giant)]
    keys, key_of_step = _keys_for(ctx, ref, L, steps, 2400)
    ev = ckh.Evaluator(ctx, None, keys)
    ev.set_config(bsgs=bsgs)
    level = L - 1
    xs_h, xs_d, ents_h, ents_d = [], [], [], []
    masks = [[(i % 3) != 1 for i in range(16)], [(i % 4) != 2 and i < 12 for i in range(16)]]
    for t, mask in enumerate(masks):
        c0, c1 = host_uniform(ref, level, 2500 + 10 * t), host_uniform(ref, level, 2501 + 10 * t)
        xs_h.append((c0, c1))
        xs_d.append(ctx.ct_upload(np.stack([c0, c1]), level, DELTA))
        eh = [host_uniform(ref, level, 2600 + 50 * t + i) if m else None for i, m in enumerate(mask)]
        ents_h.append(eh)
        ents_d.append([ctx.upload(w, level) if w is not None else None for w in eh])
    ev.reset_counter()
    got = ev.matvec_plain_sum(xs_d, ents_d, DELTA, baby, giant, stride, offset).download()
    want = oracle_matvec_sum(ref, xs_h, level, ents_h, baby, giant, stride, offset, bsgs, key_of_step, L)
    assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
    assert ev.counter().ct_pt_mults == sum(map(sum, masks))
    with pytest.raises(ckh.InvalidParams):  # a term without any entry
        ev.matvec_plain_sum(xs_d, [ents_d[0], [None] * 16], DELTA, baby, giant, stride, offset)
    with pytest.raises(ckh.ScaleMismatch):
        other = ctx.ct_upload(np.stack(xs_h[1]), level, 2 * DELTA)
        ev.matvec_plain_sum([xs_d[0], other], ents_d, DELTA, baby, giant, stride, offset)
    with pytest.raises(ckh.InvalidParams):
        ev.matvec_plain_sum([], [], DELTA, baby, giant, stride, offset)


def test_add_plain_mul_plain_match_oracle(env):
    """add_plain / mul_plain (evaluator.rs:287-308): c0 + p with a's scale (p's scale within 2^-10),
    every component times p with the product of the scales, one ct_pt_mults per mul_plain; the
    checks in the reference's order (basis, level, scale), failures counting nothing."""
    ctx, ref = env
    L = len(ref.q) - 1
    ev = ckh.Evaluator(ctx)
    for level in (L, 1, 0):
        for ncomp in (2, 3):
            ct, h = _ct(ctx, ref, level, 900 + 10 * level + ncomp, ncomp)
            w = host_uniform(ref, level, 950 + level)
            pt = ctx.upload(w, level)
            ev.reset_counter()
            s = ev.add_plain(ct, pt, DELTA * (1 + 2.0 ** -11))
            assert s.scale == DELTA and s.level == level and s.size() == ncomp
            d = s.download()
            assert (d[0] == ref.poly_add(h[0], w, level)).all()
            for c in range(1, ncomp):
                assert (d[c] == h[c]).all()
            m = ev.mul_plain(ct, pt, 3.0)
            assert m.scale == DELTA * 3.0 and m.level == level and m.size() == ncomp
            d = m.download()
            for c in range(ncomp):
                assert (d[c] == ref.poly_mul(h[c], w, level)).all(), (level, ncomp, c)
            cnt = ev.counter()
            assert cnt.ct_pt_mults == 1 and cnt.rescales == 0 and cnt.rotations == 0
    ct, _ = _ct(ctx, ref, L, 990)
    pt = ctx.upload(host_uniform(ref, L, 991), L)
    ev.reset_counter()
    with pytest.raises(ckh.ScaleMismatch):
        ev.add_plain(ct, pt, DELTA * (1 + 2.0 ** -9))
    low = ctx.upload(host_uniform(ref, L - 1, 992), L - 1)
    with pytest.raises(ckh.LevelMismatch):  # the level check comes before the scale check
        ev.add_plain(ct, low, 5.0)
    with pytest.raises(ckh.LevelMismatch):
        ev.mul_plain(ct, low, DELTA)
    with pytest.raises(ckh.InvalidParams):  # a Q*P polynomial is not over the data basis
        ev.mul_plain(ct, ctx.uniform(L, seed=993, p_limbs=len(ref.p)), DELTA)
    ctx2 = ckh.Context(ctx.log_n, ref.q, ref.p, ctx.dnum)
    foreign = ctx2.upload(host_uniform(ref, L, 994), L)
    with pytest.raises(ckh.InvalidParams):  # plaintext of another context, before the level check
        ev.add_plain(ct, foreign, DELTA)
    with pytest.raises(ckh.InvalidParams):
        ev.mul_plain(ct, foreign, DELTA)
    assert ev.counter().ct_pt_mults == 0
    del foreign
    ctx2.close()


def test_mod_switch_to_level_and_relabel(env):
    """mod_switch_to_level (evaluator.rs:513-525): the first level + 1 limbs of every component,
    scale unchanged, a copy at the same level, InvalidParams upwards; the result outlives its
    input and feeds later operations like an uploaded one. relabel (ops.rs:475-477): the same
    words under a new scale, which must be positive and finite (ciphertext.rs:40-60)."""
    ctx, ref = env
    L = len(ref.q) - 1
    ev = ckh.Evaluator(ctx)
    ct, h = _ct(ctx, ref, L, 1000, 3)
    for level in (L, L - 1, 1, 0):
        s = ev.mod_switch_to_level(ct, level)
        assert s.level == level and s.scale == DELTA and s.size() == 3
        d = s.download()
        for c in range(3):
            assert (d[c] == h[c][: level + 1]).all()
    with pytest.raises(ckh.InvalidParams):
        ev.mod_switch_to_level(ev.mod_switch_to_level(ct, 1), 2)
    two, h2 = _ct(ctx, ref, L, 1010)
    s = ev.mod_switch_to_level(two, L - 1)
    del two  # the switched ciphertext keeps its words
    r = ev.rescale(s)
    want = ref.rescale_ct(np.stack([x[:L] for x in h2]), L - 1)
    d = r.download()
    for c in range(2):
        assert (d[c] == want[c]).all()
    ev.reset_counter()
    for scale in (1.0, DELTA * 7.5, 2.0 ** -30):
        x = ev.relabel(ct, scale)
        assert x.scale == scale and x.level == L
        d = x.download()
        for c in range(3):
            assert (d[c] == h[c]).all()
    for bad in (0.0, -1.0, float("nan"), float("inf"), -float("inf")):
        with pytest.raises(ckh.InvalidParams):
            ev.relabel(ct, bad)
    cnt = ev.counter()
    assert cnt.ct_pt_mults == 0 and cnt.rescales == 0


def _fail_each_launch(ctx, ev, op):
    """Runs op with its k-th launch failing, k = 0, 1, ..., until it completes (the reference's
    failing backend, crates/ckks-train-cuda/tests/cases/failing.rs:34-59, fails the k-th
    primitive). Each failure must be a Device error that disarms the hook, leaves the operation
    counter unchanged and leaks no device memory. Returns the completed result and how many
    launches failed before it."""
    ctx.synchronize()
    live = ctx.pool_used_memory()
    k = 0
    while True:
        ev.reset_counter()
        ctx.set_tuning("fail_launch_after", k)
        try:
            got = op()
        except ckh.Device as e:
            assert "injected launch failure" in str(e)
            assert ctx.tuning("fail_launch_after") == -1
            assert all(v == 0 for v in ev.counter().as_dict().values()), k
            ctx.synchronize()
            assert ctx.pool_used_memory() == live, k
            k += 1
            assert k < 400
            continue
        ctx.set_tuning("fail_launch_after", -1)
        return got, k


def test_failure_at_every_launch_of_composite_operations():
    """A device failure at any launch of a composite operation (BSGS product with encrypted
    diagonals, hoisted rotations, multiply, rescale, rotate, a batch of key switches) is a Device
    error with nothing counted and nothing leaked; the operation then completes with the words
    it gave before (composite counts are success-only, src/matvec.rs `counted`)."""
    ctx, ref = make(C4_SHAPE_SMALL_N)
    L = len(ref.q) - 1
    baby, giant, stride, offset = 4, 2, 1, -1
    steps = [k2 * stride for k2 in range(1, baby)] + [(offset + baby * k1) * stride for k1 in range(giant)]
    keys, _ = _keys_for(ctx, ref, L, steps, 1800)
    relin = ctx.ksk_uniform(L, seed=1850)
    ev = ckh.Evaluator(ctx, relin, keys)
    x, ents = _enc_job(ctx, ref, L - 1, baby, giant, 1900, missing=(2,))
    a, _ = _ct(ctx, ref, L, 2000)
    b, _ = _ct(ctx, ref, L, 2010)
    ab = ev.mul(a, b)
    ops = {
        "matvec_enc": lambda: [ev.matvec_enc(x, ents, baby, giant, stride, offset)],
        "hoisted_rotate": lambda: ev.hoisted_rotate(a, steps[:3]),
        "mul": lambda: [ev.mul(a, b)],
        "rescale": lambda: [ev.rescale(ab)],
        "rotate": lambda: [ev.rotate(a, steps[0])],
    }
    for name, op in ops.items():
        print("op", name, flush=True)
        want = [o.download() for o in op()]
        ev.reset_counter()
        op()
        counts = ev.counter().as_dict()
        got, failed = _fail_each_launch(ctx, ev, op)
        assert failed >= 1, name
        assert ev.counter().as_dict() == counts, name
        assert all((g.download() == w).all() for g, w in zip(got, want)), name
    # the batched key switch: one program, one launch
    ds = [ctx.uniform(L - i, seed=2300 + i) for i in range(3)]
    pairs = ctx.key_switch_batch(ds, relin)
    want = [(p0.download(), p1.download()) for p0, p1 in pairs]
    ctx.set_tuning("fail_launch_after", 0)
    with pytest.raises(ckh.Device):
        ctx.key_switch_batch(ds, relin)
    assert ctx.tuning("fail_launch_after") == -1
    again = ctx.key_switch_batch(ds, relin)
    assert all((p0.download() == w0).all() and (p1.download() == w1).all() for (p0, p1), (w0, w1) in zip(again, want))


def test_upload_failure_is_a_device_error_and_leaks_nothing():
    """fail_upload makes every upload fail (failing.rs:46-49,61-67): polynomial, ciphertext and
    switching-key uploads raise Device and hold no device memory; uploads work once it is off."""
    ctx, ref = make(SMALL)
    L = len(ref.q) - 1
    h = host_uniform(ref, L, 5)
    key = ref.uniform_key(L, 7)
    ctx.synchronize()
    live = ctx.pool_used_memory()
    ctx.set_tuning("fail_upload", 1)
    assert ctx.tuning("fail_upload") == 1
    for up in (lambda: ctx.upload(h, L), lambda: ctx.ct_upload(np.stack([h, h]), L, DELTA),
               lambda: ctx.ksk_upload(key, L)):
        with pytest.raises(ckh.Device, match="injected upload failure"):
            up()
    ctx.synchronize()
    assert ctx.pool_used_memory() == live
    ctx.set_tuning("fail_upload", 0)
    assert (ctx.upload(h, L).download() == h).all()
    assert (ctx.ct_upload(np.stack([h, h]), L, DELTA).download()[1] == h).all()
    ev = ckh.Evaluator(ctx, ctx.ksk_upload(key, L))
    assert ev.has_relin_key()
