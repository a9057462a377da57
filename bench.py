"""Benchmark: bit-exact hybrid key switches per second at N = 2^16, batched vs plain.

Workload (one "step"): 256 ciphertext polynomials at N = 2^16 with limb counts drawn
uniformly from 4..24 (seeded), 24 x 50-bit q primes, 6 x 60-bit special primes, dnum 4, all
switched under one shared key in batched launches (ckh_key_switch_batch). Inputs and the key
are generated on the device before timing (resident in HBM).

Also reported: the plain path (the same 256 switches, one ciphertext per launch sequence),
the p50/p99 latency of one key switch at the full level (24 limbs), the HIP-event timing of
the dominant kernel against the HBM roofline, and the CPU oracle on a bounded sample of the
same workload. The batched results are checked word for word against the plain path, and one
switch against the CPU oracle.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("ckks-train_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

LOG_N = 16
N = 1 << LOG_N
Q_BITS = [50] * 24
P_BITS = [60] * 6
DNUM = 4
N_CT = 256
SEED = 20261202
HBM_PEAK_GBS = 8000.0


def workload_levels():
    rng = np.random.default_rng(SEED)
    limbs = rng.integers(4, 25, size=N_CT)  # limb counts 4..24
    return [int(x) - 1 for x in limbs]


def ks_bytes(level, k=6, alpha=6):
    """Algorithmic HBM bytes of one key switch: read d, read the key parts of its digits,
    write kb and ka (8N [L + 2 d_L (L + k) + 2L])."""
    L = level + 1
    d = -(-L // alpha)
    return 8 * N * (L + 2 * d * (L + k) + 2 * L)


class HipEvents:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        ms = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value


def cpu_threads():
    """(threads used, nproc, CPUs this process may run on): the box's CPU share is given by
    OMP_NUM_THREADS where set (16 per GPU), else every CPU of the affinity mask."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else affinity
    return min(threads, affinity), nproc, affinity


def cpu_baseline(q, p, levels, budget_s=8.0):
    """The CPU oracle (C restatement of the ckks-train CPU path, OpenMP over digits and limbs as
    the reference's rayon feature) on the first ciphertexts of the same workload: the sample is
    fixed by a time budget at the host's thread share, then the same sample is timed again at one
    thread."""
    import oracle as O

    threads, nproc, affinity = cpu_threads()
    ref = O.Params(LOG_N, q=q, p=p, dnum=DNUM)
    key = ref.uniform_key(23, SEED + 1)
    # inputs generated outside the timed loops (more than the budget can use)
    inputs = [O.uniform_poly(q[: lv + 1], N, SEED + 100 + i).ravel() for i, lv in enumerate(levels[:64])]

    def run(nthreads, count=None):
        O.set_threads(nthreads)
        done, t0 = 0, time.perf_counter()
        for i, lv in enumerate(levels[: len(inputs)]):
            if count is not None and done >= count:
                break
            ref.key_switch(inputs[i], lv, key, 23)
            done += 1
            if count is None and done >= 2 and time.perf_counter() - t0 > budget_s:
                break
        return done, time.perf_counter() - t0

    done, el = run(threads)
    _, el1 = run(1, done)
    O.set_threads(threads)
    return {"value": done / el, "unit": "key_switches/s", "cores": threads, "kind": "port",
            "sample": f"first {done} ciphertexts of the workload (levels {levels[:done]}), oracle "
                      f"key_switch with {threads} OpenMP threads, {el:.1f} s",
            "value_1_thread": done / el1,
            "sample_1_thread": f"the same {done} ciphertexts at 1 thread, {el1:.1f} s",
            "host_nproc": nproc, "host_affinity_cpus": affinity,
            "threads_note": "threads = the box's CPU share for one GPU (OMP_NUM_THREADS) within the affinity mask"}


def served_stream(ckh, n_ops=2000, rates=(1000.0, 2000.0, 4000.0, 6000.0, 8000.0), seed=SEED + 5):
    """The served mixed stream: Poisson arrivals (each of `rates` ops/s, n_ops per rate) over
    N = 2^16 (24 + 6 primes, dnum 4) and N = 2^15 (20 + 4 primes, dnum 5); 50% relinearize, 30%
    hoisted 16-rotation batch, 20% mul + relinearize + rescale; limbs U{4..24} (at most 20 at
    2^15); 64 rotation keys and a relinearization key resident per parameter set; 200 GB
    admission budget. One warm-up trace first (its figures are the cold start). Per rate the same
    arrival trace runs through the plain path (one op at a time, one stream), and every served
    result is compared with the plain one by device digest."""
    sets = [(16, [50] * 24, [60] * 6, 4), (15, [50] * 20, [60] * 4, 5)]
    envs = []
    for i, (log_n, qb, pb, dnum) in enumerate(sets):
        q, p = ckh.params_primes(log_n, qb, pb)
        ctx = ckh.Context(log_n, q, p, dnum, n_streams=1)
        L = len(q) - 1
        relin = ctx.ksk_uniform(L, seed=seed + 1000 * i)
        rot = [ctx.ksk_uniform(L, seed=seed + 1000 * i + s, galois=ckh.galois_elt(log_n, s)) for s in range(1, 65)]
        envs.append((ctx, L, ckh.Evaluator(ctx, relin, rot)))
    warm = _trace(ckh, envs, 400, max(rates), seed + 999)
    for op in warm[:6]:  # both contexts' kernels loaded
        _run_plain(ckh, envs, op)
    cold = _serve(ckh, envs, warm)[1]
    del warm
    out = {"cold_start": {"ops": 400, "arrival_rate_per_s": max(rates), "ops_per_s": cold.ops_per_s,
                          "p50_ms": cold.p50_ms, "p99_ms": cold.p99_ms,
                          "note": "first trace on a fresh server, pools and plans"}}
    for k, rate in enumerate(rates):
        out[f"at_{int(rate)}_per_s"] = _served_at(ckh, envs, n_ops, rate, seed + 7 * k)
        r = out[f"at_{int(rate)}_per_s"]
        print(f"# served at {rate:.0f}/s: {r['served']['ops_per_s']:.0f} ops/s p50 {r['served']['p50_ms']:.2f} "
              f"p99 {r['served']['p99_ms']:.2f} ms; plain {r['plain']['ops_per_s']:.0f} ops/s; "
              f"identical {r['identical_alone_and_in_mix']}", file=sys.stderr, flush=True)
    return out


def _trace(ckh, envs, n_ops, rate, seed):
    """A seeded arrival trace: (kind, parameter set, a, b, steps, arrival time) per op."""
    rng = np.random.default_rng(seed)
    ops, t = [], 0.0
    for i in range(n_ops):
        t += rng.exponential(1.0 / rate)
        e = int(rng.integers(0, 2))
        ctx, L, ev = envs[e]
        level = min(int(rng.integers(3, 24)), L)
        u = rng.random()

        def ct(ncomp, base):
            return ctx.ct_from_polys([ctx.uniform(level, seed=base + j) for j in range(ncomp)], 2.0 ** 40)

        if u < 0.5:
            ops.append((ckh.OP_RELINEARIZE, e, ct(3, 10 * i), None, [], t))
        elif u < 0.8:
            steps = sorted(int(s) for s in rng.choice(np.arange(1, 65), size=16, replace=False))
            ops.append((ckh.OP_HOISTED_ROTATE, e, ct(2, 10 * i), None, steps, t))
        else:
            lv = max(level, 1)
            ops.append((ckh.OP_MUL_RELIN_RESCALE, e,
                        ctx.ct_from_polys([ctx.uniform(lv, seed=10 * i + j) for j in range(2)], 2.0 ** 40),
                        ctx.ct_from_polys([ctx.uniform(lv, seed=10 * i + 5 + j) for j in range(2)], 2.0 ** 40), [], t))
    for ctx, _, _ in envs:
        ctx.synchronize()
    return ops


def _run_plain(ckh, envs, op):
    kind, e, a, b, steps, _ = op
    ev = envs[e][2]
    if kind == ckh.OP_RELINEARIZE:
        return [ev.relinearize(a)]
    if kind == ckh.OP_HOISTED_ROTATE:
        return ev.hoisted_rotate(a, steps)
    return [ev.rescale(ev.relinearize(ev.mul_no_relin(a, b)))]


def _serve(ckh, envs, ops):
    """Runs a trace through a fresh server at its arrival times; results are polled in order and
    reduced to (status, digests) at once, so outputs do not pile up. Returns (results, stats)."""
    srv = ckh.Server(hbm_budget_bytes=200 * 10 ** 9, max_batch=64)
    t0 = srv.now() + 0.02
    tickets = [srv.submit(k, envs[e][2], a, b, steps, arrival_s=t0 + ta) for (k, e, a, b, steps, ta) in ops]
    served = []
    for tk in tickets:
        status, _, outs, _ = srv.poll(tk)
        served.append((status, [o.digests() for o in outs]))
        del outs
    st = srv.stats()
    srv.close()
    return served, st


def _served_at(ckh, envs, n_ops, rate, seed):
    """One arrival trace at `rate` ops/s through the server and through the plain path."""
    ops = _trace(ckh, envs, n_ops, rate, seed)
    served, st = _serve(ckh, envs, ops)

    # plain, same arrivals, one op at a time
    lat, plain = [], []
    start = time.perf_counter() + 0.02
    for op in ops:
        arrival = start + op[5]
        while time.perf_counter() < arrival:
            pass
        out = _run_plain(ckh, envs, op)
        envs[op[1]][0].synchronize()
        lat.append((time.perf_counter() - arrival) * 1e3)
        plain.append([o.digests() for o in out])
        del out
    makespan = time.perf_counter() - start
    same = [s[0] == 0 and s[1] == p for s, p in zip(served, plain)]
    lat = np.array(lat)
    mix = {k: sum(1 for o in ops if o[0] == k) for k in (0, 1, 2)}
    del ops
    return {
        "ops": n_ops, "arrival_rate_per_s": rate,
        "mix": {"relinearize": mix[0], "hoisted_16": mix[1], "mul_relin_rescale": mix[2]},
        "served": {"ops_per_s": st.ops_per_s, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms, "mean_ms": st.mean_ms,
                   "max_ms": st.max_ms, "batches": st.batches, "peak_admitted_GB": st.peak_bytes / 1e9,
                   "failed": st.failed, "scheduler_launch_ms_total": st.launch_ms,
                   "scheduler_launch_ms_max": st.launch_max_ms},
        "plain": {"ops_per_s": n_ops / makespan, "p50_ms": float(np.percentile(lat, 50)),
                  "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean())},
        "identical_alone_and_in_mix": bool(all(same)),
        "mismatched_ops": [i for i, ok in enumerate(same) if not ok][:8],
    }


def config_table(ckh, O, ev):
    """The per-configuration figures (one JSON object): C1 NTT round trip at N = 2^13 (4 x 50-bit),
    C2 one key switch at N = 2^16 and 24 limbs, C4 the hoisted BSGS product at N = 2^15 (20 + 4
    primes, dnum 5, 32 x 32 blocks, two diagonals per block), each on the GPU and with the CPU
    oracle (1 thread and the host's threads, at most 16), and whether the GPU result equals the
    oracle's word for word."""
    from gpu_cases import C4, c4_entries_mask, host_uniform, make, oracle_matvec
    threads = min(16, os.cpu_count() or 1)
    out = {}

    def cpu_time(fn, nthreads, reps=1):
        O.set_threads(nthreads)
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        O.set_threads(threads)
        return (time.perf_counter() - t0) / reps, r

    # C1: NTT round trip, 4 limbs at N = 2^13
    n1 = 1 << 13
    q1 = O.ntt_primes(50, n1, 4)
    ctx1 = ckh.Context(13, q1, O.ntt_primes(60, n1, 1), 1)
    a = O.uniform_poly(q1, n1, 77)
    one = [ctx1.upload(a, 3)]
    batch = [ctx1.uniform(3, seed=200 + i) for i in range(256)]
    ctx1.synchronize()
    fwd1 = ctx1.time_ntt(one, forward=True, reps=50)
    inv1 = ctx1.time_ntt(one, forward=False, reps=50)
    fwdb = ctx1.time_ntt(batch, forward=True, reps=20)
    invb = ctx1.time_ntt(batch, forward=False, reps=20)
    p1 = ctx1.upload(a, 3, ntt=False)
    f1 = p1.ntt()
    rt = f1.intt().download()
    want_f = O.ntt_forward(n1, q1, a.ravel().copy())
    exact1 = bool((f1.download().ravel() == want_f.ravel()).all() and (rt == a).all())
    r1 = O.Params(13, q=q1, p=O.ntt_primes(60, n1, 1), dnum=1)
    assert (r1.ntt(r1.ntt(a, 4, True), 4, False) == a).all()
    c1_1, _ = cpu_time(lambda: r1.ntt(r1.ntt(a, 4, True), 4, False), 1, reps=20)
    c1_n, _ = cpu_time(lambda: r1.ntt(r1.ntt(a, 4, True), 4, False), threads, reps=20)
    out["C1_ntt_round_trip_n8192_4limbs"] = {
        "gpu_ms": fwd1 + inv1, "gpu_batch256_us_per_poly": (fwdb + invb) * 1e3 / 256,
        "cpu_1_thread_ms": c1_1 * 1e3, "cpu_threads_ms": c1_n * 1e3, "bit_exact": exact1}
    del ctx1, one, batch, p1, f1

    # C2: one key switch at 24 limbs
    q, p = ckh.params_primes(LOG_N, Q_BITS, P_BITS)
    ctx = ckh.Context(LOG_N, q, p, DNUM)
    ref = O.Params(LOG_N, q=q, p=p, dnum=DNUM)
    key = ctx.ksk_uniform(23, seed=SEED + 1)
    hkey = ref.uniform_key(23, SEED + 1)
    d = O.uniform_poly(q, N, 4242)
    dd = ctx.upload(d, 23)
    ctx.synchronize()
    lat = []
    for i in range(60):
        t0 = time.perf_counter()
        ctx.key_switch(dd, key)
        ctx.synchronize()
        if i >= 10:
            lat.append((time.perf_counter() - t0) * 1e3)
    kb, ka = ctx.key_switch(dd, key)
    c2_n, (wb, wa) = cpu_time(lambda: ref.key_switch(d.ravel(), 23, hkey, 23), threads)
    c2_1, _ = cpu_time(lambda: ref.key_switch(d.ravel(), 23, hkey, 23), 1)
    out["C2_key_switch_n65536_24limbs"] = {
        "gpu_ms_p50": float(np.percentile(lat, 50)), "gpu_ms_p99": float(np.percentile(lat, 99)),
        "cpu_1_thread_ms": c2_1 * 1e3, "cpu_threads_ms": c2_n * 1e3,
        "bit_exact": bool((kb.download() == wb).all() and (ka.download() == wa).all())}
    del ctx, key, dd, kb, ka, hkey

    # C4: the hoisted BSGS product
    ref4, L, baby, giant, steps, seeds, c0, c1, ents_h, times, got = bsgs_c4(ckh)
    cache = {}

    def key_of_step(st):
        if st not in cache:
            cache[st] = ref4.uniform_key(L, seeds[st])
        return cache[st]

    c4_n, want = cpu_time(lambda: oracle_matvec(ref4, c0, c1, L, ents_h, baby, giant, 1, key_of_step, L), threads)
    g = got.download()
    out["C4_hoisted_bsgs_n32768_32x32"] = {
        "gpu_ms": float(np.median(times)), "cpu_threads_ms": c4_n * 1e3, "cpu_1_thread_ms": None,
        "rotations": len(steps), "bit_exact": bool((g[0] == want[0]).all() and (g[1] == want[1]).all()),
        "note": "the 1-thread oracle run of this product (62 key switches at 21 limbs) is left out for time"}
    out["cpu_threads"] = threads
    return out


def bsgs_c4(ckh, reps=5):
    """The C4 product: N = 2^15 (20 x 50-bit + 4 x 60-bit, dnum 5), 32 x 32 BSGS blocks with the
    diagonals at baby steps k1 and k1 + 16 of block k1 (64 entries), 62 rotation keys; one
    warm-up product, then reps products timed by wall clock around a synchronize."""
    from gpu_cases import C4, c4_entries_mask, host_uniform, make
    ctx4, ref4 = make(C4)
    L = len(ref4.q) - 1
    baby = giant = 32
    mask = c4_entries_mask(baby, giant)
    steps = sorted({k2 for k2 in range(1, baby)} | {baby * k1 for k1 in range(1, giant)})
    seeds = {st: 7000 + 10 * i for i, st in enumerate(steps)}
    keys = [ctx4.ksk_uniform(L, seed=seeds[st], galois=ckh.galois_elt(ref4.log_n, st)) for st in steps]
    e4 = ckh.Evaluator(ctx4, None, keys)
    c0, c1 = host_uniform(ref4, L, 7001), host_uniform(ref4, L, 7002)
    x = ctx4.ct_upload(np.stack([c0, c1]), L, 2.0 ** 40)
    ents_h = [host_uniform(ref4, L, 8000 + k1 * baby + k2) if mask[k1][k2] else None
              for k1 in range(giant) for k2 in range(baby)]
    ents_d = [ctx4.upload(w, L) if w is not None else None for w in ents_h]
    ctx4.synchronize()
    e4.matvec_plain(x, ents_d, 2.0 ** 40, baby, giant)
    ctx4.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        got = e4.matvec_plain(x, ents_d, 2.0 ** 40, baby, giant)
        ctx4.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    return ref4, L, baby, giant, steps, seeds, c0, c1, ents_h, times, got


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "ntt_forward_240limbs_n65536.md")


def measured_traffic():
    """HBM bytes per roofline launch from the committed counter summary (rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE runs of --roofline-only, calibrated with a device copy), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_peak(ctx, nbytes=2 << 30, reps=5):
    """A streaming copy kernel over 2 GiB (eight times the 256 MiB last-level cache; 16-byte loads
    and stores, grid-stride) on the library's stream, timed with HIP events: the HBM rate a
    kernel reaches (bytes read + written)."""
    ms = ctx.time_copy(nbytes, reps)
    return 2 * nbytes / (ms * 1e-3) / 1e9


def roofline_only(ctx, ev, reps=3):
    """The roofline launch alone (forward NTT over 8 polynomials x 30 limbs), then the streaming
    copy kernel over the same byte count: the copy calibrates the HBM byte counters (its
    algorithmic traffic is exactly bytes read + bytes written)."""
    tpolys = [ctx.uniform(23, seed=SEED + 9 + i, p_limbs=6, ntt=False) for i in range(8)]
    ctx.synchronize()
    ms = ctx.time_ntt(tpolys, forward=True, reps=reps)
    nbytes = 8 * N * 8 * 30
    copy_ms = ctx.time_copy(nbytes, reps)
    ctx.synchronize()
    print(json.dumps({"ntt_forward_limbs": 240, "ntt_ms_per_launch": ms, "ntt_bytes_alg": 16 * N * 240,
                      "copy_bytes": nbytes, "copy_reps": reps + 1, "copy_ms": copy_ms}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched step (warmup + steps, no autotune) and print its time (for counter runs)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="with --batched-only or --single-only: force a tuning value (repeatable)")
    ap.add_argument("--configs", action="store_true",
                    help="print the per-configuration figures (NTT round trip, one key switch, the hoisted "
                         "BSGS product; GPU and CPU oracle) as one JSON line and exit")
    ap.add_argument("--bsgs-only", action="store_true",
                    help="run only the hoisted BSGS product at N=2^15 (3 timed products, for traces)")
    ap.add_argument("--single-only", action="store_true",
                    help="run only single key switches at 24 limbs (N=2^16; 5 untimed + 40 timed calls, for traces)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the dominant-kernel launch and a device copy of the same size (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the key-switch backend shards by ciphertext with no "
              "collectives, and multi-GPU runs are not part of this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import ckh
    import oracle as O

    if args.configs:
        print(json.dumps(config_table(ckh, O, None)), flush=True)
        return
    if args.bsgs_only:
        times = bsgs_c4(ckh, reps=3)[-2]
        print(json.dumps({"bsgs_c4_ms": times}), flush=True)
        return
    q, p = ckh.params_primes(LOG_N, Q_BITS, P_BITS)
    ctx = ckh.Context(LOG_N, q, p, DNUM)
    if args.single_only:
        for kv in args.tune:
            k, v = kv.split("=")
            ctx.set_tuning(k, int(v))
        key = ctx.ksk_uniform(23, seed=SEED + 1)
        d = ctx.uniform(23, seed=SEED + 2)
        for _ in range(5):
            ctx.key_switch(d, key)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(40):
            ctx.key_switch(d, key)
        ctx.synchronize()
        print(json.dumps({"single_switch_ms_per_call": (time.perf_counter() - t0) / 40 * 1e3}), flush=True)
        return
    stream = ctx.stream_handle()
    ev = HipEvents()
    if args.roofline_only:
        roofline_only(ctx, ev)
        return
    levels = workload_levels()
    # --batched-only runs without the autotuner, at the choices it makes on this workload
    tuning = ({"ntt_chunk_fwd": 0, "ntt_chunk_inv": 0, "conv_variant": 0, "conv_tg": 24, "fuse_ip": 1,
               "conv_variant_md": 1} if args.batched_only else ctx.autotune())
    if args.batched_only:
        for kv in args.tune:
            k, v = kv.split("=")
            tuning[k] = int(v)
        for k, v in tuning.items():
            ctx.set_tuning(k, v)
    print(f"# autotune: {json.dumps(tuning)}", file=sys.stderr, flush=True)
    key = ctx.ksk_uniform(23, seed=SEED + 1)
    polys = [ctx.uniform(lv, seed=SEED + 100 + i) for i, lv in enumerate(levels)]
    ctx.synchronize()
    if not args.batched_only:
        # the key-switch schedule chosen on the workload itself (untimed)
        tuning["ks_on_workload_ms"] = ctx.autotune_ks(polys, key)
        for k in ("conv_variant", "conv_tg", "fuse_ip", "conv_variant_md", "ntt_chunk_inv"):
            tuning[k] = ctx.tuning(k)
        print(f"# autotune on the workload: {json.dumps(tuning['ks_on_workload_ms'])}", file=sys.stderr, flush=True)
    print(f"# workload: {N_CT} polys, limbs 4..24 (mean {np.mean(levels) + 1:.1f}), key {key.info()[2] / 2**20:.0f} MiB",
          file=sys.stderr, flush=True)

    # ---- batched (the measured step)
    for _ in range(args.warmup):
        out = ctx.key_switch_batch(polys, key)
        del out
    ctx.synchronize()
    e0, e1 = ev.event(), ev.event()
    t0 = time.perf_counter()
    ev.record(e0, stream)
    for _ in range(args.steps):
        out = ctx.key_switch_batch(polys, key)
    ev.record(e1, stream)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    batched_ms = ev.elapsed_ms(e0, e1)
    ms_step = batched_ms / args.steps
    batched = N_CT / (ms_step * 1e-3)
    print(f"# batched: {ms_step:.2f} ms/step ({wall / args.steps * 1e3:.2f} ms wall)", file=sys.stderr, flush=True)
    if args.batched_only:
        print(json.dumps({"batched_ms_per_step": ms_step, "key_switches_per_s": batched, "steps": args.steps}))
        return

    # ---- plain: one ciphertext per launch sequence, same switches
    for i in range(8):
        ctx.key_switch(polys[i], key)
    ctx.synchronize()
    ev.record(e0, stream)
    plain_out = [ctx.key_switch(pl, key) for pl in polys]
    ev.record(e1, stream)
    ctx.synchronize()
    plain_ms = ev.elapsed_ms(e0, e1)
    plain = N_CT / (plain_ms * 1e-3)
    print(f"# plain: {plain_ms:.2f} ms for {N_CT}", file=sys.stderr, flush=True)

    # ---- PCIe-inclusive rate on a sample of 32 polynomials: host residues uploaded, switched in
    # one batch, both outputs downloaded (wall clock; not the metric, inputs are normally resident)
    host = [(p.download(), lv) for p, lv in zip(polys[:32], levels[:32])]
    ctx.synchronize()
    t0 = time.perf_counter()
    dev = [ctx.upload(a, lv) for a, lv in host]
    res = ctx.key_switch_batch(dev, key)
    for kb, ka in res:
        kb.download()
        ka.download()
    pcie_rate = len(host) / (time.perf_counter() - t0)
    del host, dev, res

    # ---- bit-exactness: batched == plain on every 8th ciphertext; plain == oracle on one
    exact = True
    for i in range(0, N_CT, 8):
        for a, b in zip(out[i], plain_out[i]):
            exact &= bool((a.download() == b.download()).all())
    i0 = int(np.argmin(levels))
    ref = O.Params(LOG_N, q=q, p=p, dnum=DNUM)
    O.set_threads(min(16, os.cpu_count() or 1))
    wb, wa = ref.key_switch(O.uniform_poly(q[: levels[i0] + 1], N, SEED + 100 + i0).ravel(), levels[i0],
                            ref.uniform_key(23, SEED + 1), 23)
    exact &= bool((plain_out[i0][0].download() == wb).all() and (plain_out[i0][1].download() == wa).all())
    del plain_out, out

    # ---- latency of one plain key switch at 24 limbs
    top = ctx.uniform(23, seed=SEED + 7)
    lat = []
    for i in range(40):
        ctx.synchronize()
        ev.record(e0, stream)
        r = ctx.key_switch(top, key)
        ev.record(e1, stream)
        lat.append(ev.elapsed_ms(e0, e1))
        del r
    lat = np.array(lat[5:])

    # ---- dominant kernel: the forward NTT (column pass + row pass) as one batched launch over
    # 8 polynomials of 24 Q + 6 P limbs (240 limb transforms, the limb mix of a C2 ModUp),
    # timed with HIP events on the library's stream.
    tpolys = [ctx.uniform(23, seed=SEED + 9 + i, p_limbs=6, ntt=False) for i in range(8)]
    ntt_ms = ctx.time_ntt(tpolys, forward=True, reps=20)
    limbs = 8 * (24 + 6)
    ntt_bytes = 16 * N * limbs
    achieved = ntt_bytes / (ntt_ms * 1e-3) / 1e9
    del tpolys
    ctx.synchronize()
    copy_gbs = copy_peak(ctx)
    # the same launch over data primes only (240 x 50-bit limbs, the f64 path)
    qpolys = [ctx.uniform(23, seed=SEED + 19 + i, ntt=False) for i in range(10)]
    ntt_q_ms = ctx.time_ntt(qpolys, forward=True, reps=20)
    ntt_q = {"limbs": 240, "ms_per_launch": ntt_q_ms, "GBps": 16 * N * 240 / (ntt_q_ms * 1e-3) / 1e9}
    ntt_q["frac"] = ntt_q["GBps"] / HBM_PEAK_GBS
    del qpolys

    del polys, key, top
    ctx.synchronize()
    graphs = dict(zip(("captured", "replays"), ctx.graph_stats()))
    # the served stream runs on contexts of its own: this one's cached blocks go back to the driver
    ctx.trim_pool(0)
    serve = None if args.no_serve else served_stream(ckh)
    print(f"# served stream: {json.dumps(serve)}", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_baseline(q, p, levels)

    # algorithmic bytes of one batched step: every input read and both outputs written once, the
    # shared key's limbs the batch uses read once per launch (its top level's digits)
    top_l = max(levels) + 1
    d_top = -(-top_l // 6)
    step_bytes = 8 * N * (sum(3 * (lv + 1) for lv in levels) + 2 * d_top * (top_l + 6))
    line = {
        "metric": "bit-exact key switches/sec at N=2^16 (batched vs plain) + p50/p99 op latency",
        "value": batched,
        "unit": "key_switches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "batched uneven-level hybrid key switch: 256 polynomials at N=2^16, limbs U{4..24} "
                        "(seeded), 24x50-bit q + 6x60-bit p, dnum 4, one shared key",
            "n": N, "ciphertexts_per_step": N_CT, "q_bits": 50, "nq": 24, "p_bits": 60, "np": 6, "dnum": DNUM,
            "seed": SEED, "parallelism": "single GPU",
        },
        "plain_key_switches_per_s": plain,
        "batched_over_plain": batched / plain,
        "pcie_inclusive_key_switches_per_s": pcie_rate,
        "latency_ms_24_limbs": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                "mean": float(lat.mean())},
        "bit_exact": exact,
        "effective_ks_GBps": step_bytes / (ms_step * 1e-3) / 1e9,
        "step_bytes_alg": step_bytes,
        "batched_over_plain_target": 5.0,
        "roofline": {
            "kernel": "forward NTT (column pass + row pass), one launch over 240 limbs at N=2^16 "
                      "(192 x 50-bit, 48 x 60-bit)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
            "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
            "ms_per_launch": ntt_ms, "bytes_per_launch": ntt_bytes,
            "measured_copy_GBps": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs,
        },
        "ntt_forward_q_primes_only": ntt_q,
        "tuning": tuning,
        "graphs": graphs,
        "cpu_baseline": cpu,
        "served_stream": serve,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
