"""CoeffToSlot / SlotToCoeff (crates/ckks-train-ckks/src/bootstrap/linear.rs:776-945) on the
device against the oracle composition of the same operations, with the reference's checks,
error variants and counts."""
import numpy as np
import pytest

import ckh
from gpu_cases import (C2, C2_SHAPE_N4096, C4_SHAPE_SMALL_N, bsgs_rotation_count, host_uniform, make,
                       oracle_coeff_to_slot, oracle_slot_to_coeff)

pytestmark = pytest.mark.gpu

DELTA = 2.0 ** 40


def say(*a):
    print("case:", *a, flush=True)


def _setup(case, shapes, seed):
    """Context, oracle, evaluator with every rotation key the shapes use and the conjugation
    key, and the host key lookup."""
    ctx, ref = make(case)
    L = len(ref.q) - 1
    slots = ref.n // 2
    steps = set()
    for baby, giant, stride, offset in shapes:
        steps.update(k2 * stride for k2 in range(1, baby))
        steps.update((offset + baby * k1) * stride for k1 in range(giant))
    norm = sorted({s % slots for s in steps if s % slots})
    keys, host = [], {}
    for i, st in enumerate(norm):
        keys.append(ctx.ksk_uniform(L, seed=seed + 10 * i, galois=ckh.galois_elt(ref.log_n, st)))
        host[st] = seed + 10 * i
    conj_seed = seed + 9999
    keys.append(ctx.ksk_uniform(L, seed=conj_seed, galois=ckh.conjugation_elt(ref.log_n)))
    cache = {}

    def key_of_step(st):
        st %= slots
        if st not in cache:
            cache[st] = ref.uniform_key(L, host[st])
        return cache[st]

    ev = ckh.Evaluator(ctx, None, keys)
    return ctx, ref, L, ev, key_of_step, ref.uniform_key(L, conj_seed), keys


def _diags(ctx, ref, level, shape, mask, seed):
    baby, giant = shape[0], shape[1]
    assert len(mask) == baby * giant
    h = [host_uniform(ref, level, seed + i) if m else None for i, m in enumerate(mask)]
    d = [ctx.upload(w, level) if w is not None else None for w in h]
    return h, d


def _plan(ev, levels, s_in, target=DELTA):
    """(stage scale, diagonal scale) per stage: the diagonals take the planned input scale to
    the stage's exact pre-rescale label."""
    out = []
    for lv in levels:
        label = ev.exact_product_scale(lv, target)
        out.append((target, label / s_in))
        s_in = target
    return out


@pytest.mark.parametrize("case", [C4_SHAPE_SMALL_N, C2_SHAPE_N4096, C2], ids=lambda c: c[0])
def test_coeff_to_slot_matches_oracle(case):
    say("coeff_to_slot", case[0])
    shapes = [(2, 2, 1, 0), (4, 2, 2, -3)]
    ctx, ref, L, ev, key_of_step, conj_key, _ = _setup(case, shapes, 7100)
    lv0 = L - 1
    levels = [lv0, lv0 - 1]
    masks = [[True, False, True, True], [i % 3 != 2 for i in range(8)]]
    masks2 = [None, [i % 2 == 0 or i == 5 for i in range(8)]]
    plan = _plan(ev, levels, DELTA)
    stages, host_stages = [], []
    for k, (lv, shape) in enumerate(zip(levels, shapes)):
        h1, d1 = _diags(ctx, ref, lv, shape, masks[k], 7300 + 100 * k)
        h2, d2 = (None, None) if masks2[k] is None else _diags(ctx, ref, lv, shape, masks2[k], 7350 + 100 * k)
        stages.append(ckh.LinearStage(ckh.COEFF_TO_SLOT, lv, plan[k][0], d1, plan[k][1], shape, second=d2))
        host_stages.append((lv, h1, h2, shape))
    c0, c1 = host_uniform(ref, lv0, 7001), host_uniform(ref, lv0, 7002)
    ct = ctx.ct_upload(np.stack([c0, c1]), lv0, DELTA)
    ev.reset_counter()
    re, im = ev.coeff_to_slot(ct, stages)
    want_re, want_im = oracle_coeff_to_slot(ref, c0, c1, host_stages, key_of_step, conj_key, L)
    got_re, got_im = re.download(), im.download()
    for g, w in ((got_re, want_re), (got_im, want_im)):
        assert (g[0] == w[0]).all() and (g[1] == w[1]).all()
    assert re.level == im.level == lv0 - 2
    assert re.scale == im.scale == plan[-1][0]
    slots = ref.n // 2
    rot = bsgs_rotation_count([masks[0]], *shapes[0], slots) + bsgs_rotation_count([masks[1], masks2[1]], *shapes[1], slots)
    c = ev.counter()
    assert c.rotations == rot + 2  # and the two conjugations
    assert c.rescales == 2
    assert c.ct_pt_mults == sum(masks[0]) + sum(masks[1]) + sum(masks2[1])


@pytest.mark.parametrize("case", [C4_SHAPE_SMALL_N, C2_SHAPE_N4096], ids=lambda c: c[0])
def test_slot_to_coeff_matches_oracle(case):
    say("slot_to_coeff", case[0])
    shapes = [(4, 2, 1, -1), (2, 4, 4, 0)]
    ctx, ref, L, ev, key_of_step, _, _ = _setup(case, shapes, 8100)
    lv0 = L - 2
    levels = [lv0, lv0 - 1]
    masks = [[i != 3 for i in range(8)], [i % 2 == 1 for i in range(8)]]
    masks2 = [[i % 3 == 0 for i in range(8)], None]
    plan = _plan(ev, levels, DELTA)
    stages, host_stages = [], []
    for k, (lv, shape) in enumerate(zip(levels, shapes)):
        h1, d1 = _diags(ctx, ref, lv, shape, masks[k], 8300 + 100 * k)
        h2, d2 = (None, None) if masks2[k] is None else _diags(ctx, ref, lv, shape, masks2[k], 8350 + 100 * k)
        stages.append(ckh.LinearStage(ckh.SLOT_TO_COEFF, lv, plan[k][0], d1, plan[k][1], shape, second=d2))
        host_stages.append((lv, h1, h2, shape))
    hr = (host_uniform(ref, lv0, 8001), host_uniform(ref, lv0, 8002))
    hi = (host_uniform(ref, lv0, 8003), host_uniform(ref, lv0, 8004))
    re = ctx.ct_upload(np.stack(hr), lv0, DELTA)
    im = ctx.ct_upload(np.stack(hi), lv0, DELTA)
    ev.reset_counter()
    out = ev.slot_to_coeff(re, im, stages)
    want = oracle_slot_to_coeff(ref, hr, hi, host_stages, key_of_step, L)
    got = out.download()
    assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
    assert out.level == lv0 - 2 and out.scale == plan[-1][0]
    slots = ref.n // 2
    c = ev.counter()
    assert c.rotations == (bsgs_rotation_count([masks[0], masks2[0]], *shapes[0], slots)
                           + bsgs_rotation_count([masks[1]], *shapes[1], slots))
    assert c.rescales == 2
    assert c.ct_pt_mults == sum(masks[0]) + sum(masks2[0]) + sum(masks[1])


def test_linear_transform_errors_before_any_work():
    """The checks of coeff_to_slot / slot_to_coeff / check_stages, each before any work: the
    counter stays at zero."""
    say("linear errors")
    case = C4_SHAPE_SMALL_N
    shape = (2, 2, 1, 0)
    ctx, ref, L, ev, _, _, keys = _setup(case, [shape], 9100)
    lv = L - 1
    (s1, p1), = _plan(ev, [lv], DELTA)
    (s2, p2), = _plan(ev, [lv - 1], s1)
    _, d = _diags(ctx, ref, lv, shape, [True] * 4, 9300)
    _, db = _diags(ctx, ref, lv - 1, shape, [True] * 4, 9400)
    ct = ctx.ct_upload(np.stack([host_uniform(ref, lv, 9001), host_uniform(ref, lv, 9002)]), lv, DELTA)
    C2S, S2C = ckh.COEFF_TO_SLOT, ckh.SLOT_TO_COEFF

    def st(direction=C2S, level=lv, scale=s1, diags=d, pt=p1, second=None, second_pt=None, sh=shape, sh2=None):
        return ckh.LinearStage(direction, level, scale, diags, pt, sh, second=second, second_pt_scale=second_pt,
                               second_shape=sh2)

    good_last = st(second=d)
    ev.reset_counter()
    cases = [
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [])),
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st(S2C, second=d)])),
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st()])),  # the last stage needs the conjugate
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st(second=d), st(level=lv - 1, scale=s2, diags=db, pt=p2,
                                                                           second=db)])),
        (ckh.LevelMismatch, lambda: ev.coeff_to_slot(ct, [st(level=lv - 1, diags=db, second=db)])),
        (ckh.ScaleMismatch, lambda: ev.coeff_to_slot(ct, [st(pt=p1 * 2, second=d, second_pt=p1 * 2)])),
        (ckh.ScaleMismatch, lambda: ev.coeff_to_slot(ct, [st(second=d, second_pt=p1 * 2)])),
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st(diags=[None] * 4, second=d)])),
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st(second=[d[0], d[1]], sh2=(1, 2, 1, 0))])),
        (ckh.InvalidParams, lambda: ev.coeff_to_slot(ct, [st(sh=(2, 2, 3, 0), second=d)])),  # stride
        (ckh.InvalidParams, lambda: ev.slot_to_coeff(ct, ct, [good_last])),  # direction
        (ckh.InvalidParams, lambda: ev.slot_to_coeff(ct, ct, [st(S2C)])),  # the first stage takes two inputs
        (ckh.InvalidParams, lambda: ev.slot_to_coeff(ct, ct, [])),
    ]
    ct_low = ctx.ct_upload(np.stack([host_uniform(ref, lv - 1, 9003), host_uniform(ref, lv - 1, 9004)]), lv - 1, DELTA)
    ct_other = ctx.ct_upload(np.stack([host_uniform(ref, lv, 9005), host_uniform(ref, lv, 9006)]), lv, 2 * DELTA)
    cases += [
        (ckh.LevelMismatch, lambda: ev.slot_to_coeff(ct, ct_low, [st(S2C, second=d)])),
        (ckh.ScaleMismatch, lambda: ev.slot_to_coeff(ct, ct_other, [st(S2C, second=d)])),
        (ckh.ScaleMismatch, lambda: ev.slot_to_coeff(ct_other, ct_other, [st(S2C, second=d)])),
    ]
    for i, (err, fn) in enumerate(cases):
        with pytest.raises(err):
            fn()
    # a missing rotation key, and a missing conjugation key
    no_rot = ckh.Evaluator(ctx, None, [keys[-1]])
    with pytest.raises(ckh.InvalidParams):
        no_rot.coeff_to_slot(ct, [good_last])
    no_conj = ckh.Evaluator(ctx, None, keys[:-1])
    with pytest.raises(ckh.InvalidParams):
        no_conj.coeff_to_slot(ct, [good_last])
    c = ev.counter()
    assert (c.rotations, c.rescales, c.ct_pt_mults) == (0, 0, 0)
    # the valid one-stage transform runs
    re, im = ev.coeff_to_slot(ct, [good_last])
    assert re.level == lv - 1 and im.scale == s1
