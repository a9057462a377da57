"""The no-allocation batched key switch (ckh_key_switch_batch_static) against the allocating one
(ckh_key_switch_batch), through the C ABI: two calls in a row, the second on a smaller batch,
word for word; no device memory taken from the context and no graph captured by either call;
batches beyond the fixed workspace and overlapping outputs rejected before any work."""
import sys

import numpy as np
import pytest

import ckh
from gpu_cases import C2, C2_SHAPE_N4096, make

pytestmark = pytest.mark.gpu

WORKSPACE = 256 << 20  # CKH_STATIC_SCRATCH_BYTES


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _outputs(ctx, levels, seed):
    """Existing polynomials to be overwritten (uniform, so stale words would show)."""
    return [(ctx.uniform(lv, seed=seed + 2 * i), ctx.uniform(lv, seed=seed + 2 * i + 1)) for i, lv in enumerate(levels)]


def _check_calls(ctx, key, batches, seed):
    for call, levels in enumerate(batches):
        say("static batch call", call, "levels", levels)
        polys = [ctx.uniform(lv, seed=seed + 100 * call + i) for i, lv in enumerate(levels)]
        outs = _outputs(ctx, levels, seed + 7000 + 100 * call)
        ctx.synchronize()
        live0 = ctx.memory_info()[0]
        graphs0 = ctx.graph_stats()
        ctx.key_switch_batch_static(polys, key, outs)
        ctx.synchronize()
        assert ctx.memory_info()[0] == live0, "the static switch took device memory"
        assert ctx.graph_stats() == graphs0, "the static switch captured or replayed a graph"
        want = ctx.key_switch_batch(polys, key)
        for (gb, ga), (wb, wa) in zip(outs, want):
            assert (gb.download() == wb.download()).all()
            assert (ga.download() == wa.download()).all()


def test_static_switch_equals_batched_small_n():
    ctx, ref = make(C2_SHAPE_N4096)
    L = len(ref.q) - 1
    key = ctx.ksk_uniform(L, seed=41)
    rng = np.random.default_rng(7)
    first = [int(x) for x in rng.integers(0, L + 1, size=9)] + [L]
    second = [int(x) for x in rng.integers(0, L + 1, size=3)]
    _check_calls(ctx, key, [first, second], seed=900)
    ctx.synchronize()


def test_static_switch_equals_batched_full_size():
    """N = 2^16, 24 + 6 primes, dnum 4: a top-level switch and a 4-limb one, then one 12-limb
    switch (the workspace keeps the first call's words; nothing of them may leak)."""
    ctx, ref = make(C2)
    key = ctx.ksk_uniform(23, seed=42)
    _check_calls(ctx, key, [[23, 3], [11]], seed=1900)
    ctx.synchronize()


def test_static_switch_rejects_batches_beyond_workspace_and_overlaps():
    ctx, ref = make(C2)
    key = ctx.ksk_uniform(23, seed=43)
    levels = [23, 23]  # 2 x 132 MiB of intermediates: above the workspace
    polys = [ctx.uniform(lv, seed=2900 + i) for i, lv in enumerate(levels)]
    outs = _outputs(ctx, levels, 3900)
    before = [(b.digest(), a.digest()) for b, a in outs]
    with pytest.raises(ckh.InvalidParams):
        ctx.key_switch_batch_static(polys, key, outs)
    assert [(b.digest(), a.digest()) for b, a in outs] == before
    p = polys[0]
    with pytest.raises(ckh.InvalidParams):  # output aliasing the input
        ctx.key_switch_batch_static([p], key, [(p, outs[0][1])])
    with pytest.raises(ckh.InvalidParams):  # both outputs the same polynomial
        ctx.key_switch_batch_static([p], key, [(outs[0][0], outs[0][0])])
    with pytest.raises(ckh.InvalidParams):  # output at another level
        ctx.key_switch_batch_static([p], key, [(ctx.uniform(5, seed=1), outs[0][1])])
    ctx.synchronize()
