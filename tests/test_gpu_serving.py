"""Serving loop: continuous batching over two parameter sets must return, for every request,
exactly what the same operation returns alone; an injected failure fails only its request; a
request above the HBM budget is rejected with a Device error."""
import numpy as np
import pytest

import ckh
from gpu_cases import make

pytestmark = pytest.mark.gpu

SETS = [("c2_shape_n4096", 12, [50] * 24, [60] * 6, 4), ("c4_shape_n1024", 10, [50] * 20, [60] * 4, 5)]


def _setup(seed):
    envs = []
    for i, case in enumerate(SETS):
        ctx, ref = make(case, n_streams=3)
        L = len(ref.q) - 1
        relin = ctx.ksk_uniform(L, seed=seed + 100 * i)
        gk = [ctx.ksk_uniform(L, seed=seed + 100 * i + s, galois=ckh.galois_elt(ref.log_n, s)) for s in range(1, 17)]
        envs.append((ctx, ref, ckh.Evaluator(ctx, relin, gk)))
    return envs


def _requests(envs, n, seed, rate=4000.0):
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        e = int(rng.integers(0, len(envs)))
        ctx, ref, ev = envs[e]
        L = len(ref.q) - 1
        level = int(rng.integers(3, L + 1))
        u = rng.random()
        if u < 0.5:
            kind = ckh.OP_RELINEARIZE
            a = ctx.ct_upload(np.stack([ctx.uniform(level, seed + 10 * i + j).download() for j in range(3)]), level,
                              2.0 ** 40)
            reqs.append((kind, e, a, None, [], t))
        elif u < 0.8:
            kind = ckh.OP_HOISTED_ROTATE
            a = ctx.ct_upload(np.stack([ctx.uniform(level, seed + 10 * i + j).download() for j in range(2)]), level,
                              2.0 ** 40)
            steps = list(range(1, 17))
            reqs.append((kind, e, a, None, steps, t))
        else:
            kind = ckh.OP_MUL_RELIN_RESCALE
            level = max(level, 1)
            a = ctx.ct_upload(np.stack([ctx.uniform(level, seed + 10 * i + j).download() for j in range(2)]), level,
                              2.0 ** 40)
            b = ctx.ct_upload(np.stack([ctx.uniform(level, seed + 10 * i + 5 + j).download() for j in range(2)]), level,
                              2.0 ** 40)
            reqs.append((kind, e, a, b, [], t))
    return reqs


def _alone(envs, req):
    kind, e, a, b, steps, _ = req
    ev = envs[e][2]
    if kind == ckh.OP_RELINEARIZE:
        return [ev.relinearize(a)]
    if kind == ckh.OP_HOISTED_ROTATE:
        return ev.hoisted_rotate(a, steps)
    return [ev.rescale(ev.relinearize(ev.mul_no_relin(a, b)))]


@pytest.mark.parametrize("rate,pool", [(4000.0, 1), (800.0, 1), (4000.0, 0), (800.0, 0)])
def test_served_mix_equals_alone_and_failure_isolated(rate, pool):
    """At a high arrival rate (large batches) and a low one (many one-request batches spread
    over the streams, repeated shapes replayed from captured graphs), with the context's block
    cache and with the device's default stream-ordered pool (buffers freed while batches run on
    the four service streams and the caller's)."""
    envs = _setup(7000)
    for ctx, _, _ in envs:
        ctx.set_tuning("pool", pool)
    reqs = _requests(envs, 40, 7, rate)
    srv = ckh.Server(hbm_budget_bytes=64 << 30, max_batch=16)
    t0 = srv.now() + 0.05
    bad = 11
    tickets = [srv.submit(k, envs[e][2], a, b, steps, arrival_s=t0 + t, inject_failure=(i == bad))
               for i, (k, e, a, b, steps, t) in enumerate(reqs)]
    results = [srv.poll(t, wait=True) for t in tickets]
    st = srv.stats()
    assert st.submitted == 40 and st.completed == 39 and st.failed == 1 and st.rejected == 0
    assert st.batches >= 2 and st.p99_ms >= st.p50_ms > 0
    for i, ((status, msg, outs, lat), req) in enumerate(zip(results, reqs)):
        if i == bad:
            assert status == ckh.Device.status and "injected" in msg and outs == []
            continue
        assert status == 0, msg
        assert lat > 0
        alone = _alone(envs, req)
        assert len(outs) == len(alone)
        for o, w in zip(outs, alone):
            assert o.info() == w.info()
            assert o.digests() == w.digests(), (i, req[0])
    # spot-check one result word for word as well
    status, _, outs, _ = results[0]
    alone = _alone(envs, reqs[0])
    assert (outs[0].download() == alone[0].download()).all()
    srv.close()


def test_request_above_budget_is_rejected():
    envs = _setup(8000)
    ctx, ref, ev = envs[0]
    L = len(ref.q) - 1
    a = ctx.ct_upload(np.stack([ctx.uniform(L, 1 + j).download() for j in range(3)]), L, 2.0 ** 40)
    srv = ckh.Server(hbm_budget_bytes=1 << 20)
    t = srv.submit(ckh.OP_RELINEARIZE, ev, a)
    status, msg, outs, _ = srv.poll(t)
    assert status == ckh.Device.status and "budget" in msg and outs == []
    assert srv.stats().rejected == 1
    srv.close()


def test_caller_keeps_working_while_serving():
    """The caller's thread runs its own operations and digests on its stream while the scheduler
    runs batches (the scheduler works on streams of its own, so neither disturbs the other's
    launches or graph captures), and releases each served result as soon as it has checked it
    (a release from another stream than the allocation's)."""
    envs = _setup(9000)
    reqs = _requests(envs, 40, 9, 1500.0)
    srv = ckh.Server(hbm_budget_bytes=64 << 30, max_batch=16)
    t0 = srv.now() + 0.02
    tickets = [srv.submit(k, envs[e][2], a, b, steps, arrival_s=t0 + t) for (k, e, a, b, steps, t) in reqs]
    for i, (tk, req) in enumerate(zip(tickets, reqs)):
        alone = _alone(envs, req)  # caller-stream work while later requests are being served
        status, msg, outs, _ = srv.poll(tk, wait=True)
        assert status == 0, msg
        assert [o.digests() for o in outs] == [w.digests() for w in alone], (i, req[0])
        del outs, alone
    assert srv.stats().completed == len(reqs)
    srv.close()


def test_pool_mode_values():
    """Pool mode 1 (block cache) is the default; 0 (the device's pool) may be selected; other
    values are refused."""
    envs = _setup(9500)
    ctx = envs[0][0]
    assert ctx.tuning("pool") == 1
    with pytest.raises(ckh.InvalidParams):
        ctx.set_tuning("pool", 2)
    ctx.set_tuning("pool", 0)
    assert ctx.tuning("pool") == 0
    ctx.set_tuning("pool", 1)
    assert ctx.tuning("pool") == 1


FULL_SETS = [("c2", 16, [50] * 24, [60] * 6, 4), ("c4", 15, [50] * 20, [60] * 4, 5)]
N_ROT_KEYS = 64


def test_served_mix_full_size_64_resident_keys():
    """The served mixed stream at its stated size: N = 2^16 (24 + 6 primes, dnum 4) and N = 2^15
    (20 + 4, dnum 5), 64 rotation keys and a relinearization key resident per set, 48 Poisson
    arrivals of the 50/30/20 mix at levels 3..top. Every result equals the same operation run
    alone (by digest); three members (one of each kind, both sets) equal the oracle's composition
    word for word."""
    import oracle as O

    O.set_threads(16)
    envs, refs = [], []
    for i, case in enumerate(FULL_SETS):
        print("set", case[0], flush=True)
        ctx, ref = make(case, n_streams=2)
        L = len(ref.q) - 1
        relin = ctx.ksk_uniform(L, seed=60000 + 100 * i)
        gk = [ctx.ksk_uniform(L, seed=60000 + 100 * i + s, galois=ckh.galois_elt(case[1], s))
              for s in range(1, N_ROT_KEYS + 1)]
        envs.append((ctx, ref, ckh.Evaluator(ctx, relin, gk)))
        refs.append((relin, gk))
    rng = np.random.default_rng(61)
    reqs, t = [], 0.0
    for i in range(48):
        t += rng.exponential(1.0 / 4000.0)
        e = int(rng.integers(0, 2))
        ctx, ref, ev = envs[e]
        L = len(ref.q) - 1
        level = int(rng.integers(3, L + 1))
        u = rng.random()
        comps = lambda k, s: ctx.ct_from_polys([ctx.uniform(level, seed=s + j) for j in range(k)], 2.0 ** 40)
        if u < 0.5:
            reqs.append((ckh.OP_RELINEARIZE, e, comps(3, 70000 + 10 * i), None, [], t, level))
        elif u < 0.8:
            reqs.append((ckh.OP_HOISTED_ROTATE, e, comps(2, 70000 + 10 * i), None, list(range(1, 17)), t, level))
        else:
            reqs.append((ckh.OP_MUL_RELIN_RESCALE, e, comps(2, 70000 + 10 * i), comps(2, 70005 + 10 * i), [], t,
                         level))
    srv = ckh.Server(hbm_budget_bytes=200 * 10 ** 9, max_batch=16)
    t0 = srv.now() + 0.05
    tickets = [srv.submit(k, envs[e][2], a, b, steps, arrival_s=t0 + ta) for (k, e, a, b, steps, ta, _) in reqs]
    results = [srv.poll(tk, wait=True) for tk in tickets]
    st = srv.stats()
    assert st.completed == 48 and st.failed == 0 and st.rejected == 0
    for i, ((status, msg, outs, _), req) in enumerate(zip(results, reqs)):
        assert status == 0, msg
        alone = _alone(envs, req[:6])
        assert [o.digests() for o in outs] == [w.digests() for w in alone], (i, req[0])
    srv.close()
    # one member of each kind against the oracle composition, words downloaded from the served result
    checked = set()
    for i, (req, (_, _, outs, _)) in enumerate(zip(reqs, results)):
        kind, e, a, b, steps, _, level = req
        if kind in checked or (kind == ckh.OP_HOISTED_ROTATE and e != 1):
            continue
        checked.add(kind)
        ctx, ref, ev = envs[e]
        L = len(ref.q) - 1
        relin, gk = refs[e]
        print("oracle check", i, kind, "set", e, "level", level, flush=True)
        h = a.download()
        if kind == ckh.OP_RELINEARIZE:
            w0, w1 = ref.relinearize(h[0], h[1], h[2], level, ref.uniform_key(L, 60000 + 100 * e), L)
            d = outs[0].download()
            assert (d[0] == w0).all() and (d[1] == w1).all()
        elif kind == ckh.OP_HOISTED_ROTATE:
            elts = np.array([ckh.galois_elt(ctx.log_n, s) for s in steps], dtype=np.uint64)
            keys = [ref.uniform_key(L, 60000 + 100 * e + s) for s in steps]
            want = ref.hoisted_galois(h[0], h[1], level, elts, keys, L)
            for j, o in enumerate(outs):
                d = o.download()
                assert (d[0] == want[j][0]).all() and (d[1] == want[j][1]).all(), j
        else:
            hb = b.download()
            d3 = ref.mul_no_relin(h[0], h[1], hb[0], hb[1], level)
            r0, r1 = ref.relinearize(d3[0], d3[1], d3[2], level, ref.uniform_key(L, 60000 + 100 * e), L)
            want = ref.rescale_ct(np.stack([r0, r1]), level)
            d = outs[0].download()
            assert (d[0] == want[0]).all() and (d[1] == want[1]).all()
    assert len(checked) == 3, checked
