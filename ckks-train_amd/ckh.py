"""Python binding of libckh.so (include/ckh.h) through ctypes.

Mirrors the reference's backend surface: `Evaluator` carries the method names of the
`HeOps` trait (crates/ckks-train-ckks/src/ops.rs:198-356) and raises `HekitError` subclasses
named after the reference's error variants (crates/ckks-train-core/src/error.rs:7-21). There
is no fallback: if the HIP library is not built, loading fails with an error.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libckh.so")

u64 = C.c_uint64
u32 = C.c_uint32
P64 = C.POINTER(C.c_uint64)
VP = C.c_void_p


class HekitError(Exception):
    status = None


class InvalidParams(HekitError):
    status = 1


class ScaleMismatch(HekitError):
    status = 2


class LevelMismatch(HekitError):
    status = 3


class NoLevelsLeft(HekitError):
    status = 4


class Decode(HekitError):
    status = 5


class Device(HekitError):
    status = 6


_ERRORS = {c.status: c for c in (InvalidParams, ScaleMismatch, LevelMismatch, NoLevelsLeft, Decode, Device)}


class Params(C.Structure):
    _fields_ = [("log_n", u32), ("q", P64), ("nq", u32), ("p", P64), ("np", u32), ("dnum", u32)]


class OpCounter(C.Structure):
    _fields_ = [(f, u64) for f in ("rotations", "relins", "ct_ct_tensor", "ct_pt_mults", "rescales",
                                   "sse_switches", "bootstraps")]

    def key_switches(self):
        return self.rotations + self.relins + self.sse_switches

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", u64), ("max_batch", u32), ("max_inflight", u32)]


class Request(C.Structure):
    _fields_ = [("kind", u32), ("ev", VP), ("a", VP), ("b", VP), ("steps", C.POINTER(C.c_int32)),
                ("n_steps", u32), ("arrival_s", C.c_double), ("inject_failure", u32)]


class ServerStats(C.Structure):
    _fields_ = [(f, u64) for f in ("submitted", "completed", "failed", "rejected", "batches", "peak_bytes")] + \
               [(f, C.c_double) for f in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "makespan_s", "ops_per_s",
                                          "launch_ms", "launch_max_ms")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class BsgsShape(C.Structure):
    _fields_ = [("baby", u32), ("giant", u32), ("stride", u32), ("offset", C.c_int32)]


class LinearStageC(C.Structure):
    _fields_ = [("slot_to_coeff", C.c_int32), ("level", u32), ("scale", C.c_double), ("shape", BsgsShape),
                ("diags", C.POINTER(VP)), ("pt_scale", C.c_double), ("second_shape", BsgsShape),
                ("second", C.POINTER(VP)), ("second_pt_scale", C.c_double)]


COEFF_TO_SLOT, SLOT_TO_COEFF = 0, 1


class LinearStage:
    """One stage of CoeffToSlot / SlotToCoeff (LinearTransform,
    crates/ckks-train-ckks/src/bootstrap/linear.rs:467-502): `diags` (Poly or None, baby * giant
    of them, pre-rotated) at `level` with scale `pt_scale`, in `shape` = (baby, giant, stride,
    offset); `second` the diagonals of a two-input stage (None: one input); `scale` the exact
    scale after the stage's rescale."""

    def __init__(self, direction, level, scale, diags, pt_scale, shape, second=None, second_pt_scale=None,
                 second_shape=None):
        self.direction, self.level, self.scale = direction, level, scale
        self.diags, self.pt_scale, self.shape = list(diags), pt_scale, tuple(shape)
        self.second = None if second is None else list(second)
        self.second_pt_scale = pt_scale if second_pt_scale is None else second_pt_scale
        self.second_shape = self.shape if second_shape is None else tuple(second_shape)


OP_RELINEARIZE, OP_HOISTED_ROTATE, OP_MUL_RELIN_RESCALE = 0, 1, 2

# name: (restype, argtypes)
_SIGS = {
    "ckh_last_error": (C.c_char_p, []),
    "ckh_params_primes": (C.c_int, [u32, C.POINTER(u32), u32, C.POINTER(u32), u32, P64, P64]),
    "ckh_galois_elt": (u64, [u32, C.c_int32]),
    "ckh_conjugation_elt": (u64, [u32]),
    "ckh_device_open": (C.c_int, [C.c_int, C.POINTER(Params), u32, C.POINTER(VP)]),
    "ckh_device_close": (None, [VP]),
    "ckh_set_stream": (C.c_int, [VP, u32]),
    "ckh_stream_handle": (VP, [VP]),
    "ckh_synchronize": (C.c_int, [VP]),
    "ckh_autotune": (C.c_int, [VP, C.c_char_p, C.c_size_t]),
    "ckh_autotune_ks": (C.c_int, [VP, C.c_uint32, C.POINTER(VP), VP, C.c_char_p, C.c_size_t]),
    "ckh_set_tuning": (C.c_int, [VP, C.c_char_p, C.c_int64]),
    "ckh_memory_info": (C.c_int, [VP, P64, P64]),
    "ckh_graph_stats": (C.c_int, [VP, P64, P64]),
    "ckh_get_tuning": (C.c_int, [VP, C.c_char_p, C.POINTER(C.c_int64)]),
    "ckh_poly_upload": (C.c_int, [VP, P64, u32, u32, C.c_int, C.POINTER(VP)]),
    "ckh_poly_download": (C.c_int, [VP, VP, P64]),
    "ckh_poly_info": (C.c_int, [VP, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_int)]),
    "ckh_poly_digest": (C.c_int, [VP, VP, P64]),
    "ckh_poly_retain": (VP, [VP]),
    "ckh_poly_release": (None, [VP]),
    "ckh_poly_uniform": (C.c_int, [VP, u32, u32, C.c_int, u64, u64, C.POINTER(VP)]),
    "ckh_ntt": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_intt": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_automorphism": (C.c_int, [VP, VP, u64, C.POINTER(VP)]),
    "ckh_rescale_poly": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_time_ntt": (C.c_int, [VP, C.POINTER(VP), u32, C.c_int, u32, C.POINTER(C.c_double)]),
    "ckh_mod_down": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_time_copy": (C.c_int, [VP, u64, u32, C.POINTER(C.c_double)]),
    "ckh_ksk_upload": (C.c_int, [VP, P64, u32, u64, C.POINTER(VP)]),
    "ckh_ksk_uniform": (C.c_int, [VP, u32, u64, u64, C.POINTER(VP)]),
    "ckh_ksk_info": (C.c_int, [VP, C.POINTER(u32), P64, P64]),
    "ckh_ksk_release": (None, [VP]),
    "ckh_decompose": (C.c_int, [VP, VP, C.POINTER(VP), C.POINTER(u32)]),
    "ckh_apply_key": (C.c_int, [VP, C.POINTER(VP), u32, VP, C.POINTER(VP), C.POINTER(VP)]),
    "ckh_key_switch": (C.c_int, [VP, VP, VP, C.POINTER(VP), C.POINTER(VP)]),
    "ckh_key_switch_batch": (C.c_int, [VP, u32, C.POINTER(VP), VP, C.POINTER(VP), C.POINTER(VP)]),
    "ckh_key_switch_batch_static": (C.c_int, [VP, u32, C.POINTER(VP), VP, C.POINTER(VP), C.POINTER(VP)]),
    "ckh_ct_upload": (C.c_int, [VP, P64, u32, u32, C.c_double, C.POINTER(VP)]),
    "ckh_ct_from_polys": (C.c_int, [VP, C.POINTER(VP), u32, C.c_double, C.POINTER(VP)]),
    "ckh_ct_download": (C.c_int, [VP, VP, P64]),
    "ckh_ct_info": (C.c_int, [VP, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_double)]),
    "ckh_ct_component": (VP, [VP, u32]),
    "ckh_ct_retain": (VP, [VP]),
    "ckh_ct_release": (None, [VP]),
    "ckh_evaluator_create": (C.c_int, [VP, VP, C.POINTER(VP), u32, C.POINTER(VP)]),
    "ckh_evaluator_fork": (C.c_int, [VP, C.POINTER(VP)]),
    "ckh_evaluator_destroy": (None, [VP]),
    "ckh_counter": (C.c_int, [VP, C.POINTER(OpCounter)]),
    "ckh_reset_counter": (C.c_int, [VP]),
    "ckh_add_counts": (C.c_int, [VP, C.POINTER(OpCounter)]),
    "ckh_has_relin_key": (C.c_int, [VP]),
    "ckh_has_galois_key": (C.c_int, [VP, C.c_int32]),
    "ckh_owns": (C.c_int, [VP, VP]),
    "ckh_add": (C.c_int, [VP, VP, VP, C.POINTER(VP)]),
    "ckh_sub": (C.c_int, [VP, VP, VP, C.POINTER(VP)]),
    "ckh_negate": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_add_plain": (C.c_int, [VP, VP, VP, C.c_double, C.POINTER(VP)]),
    "ckh_mul_plain": (C.c_int, [VP, VP, VP, C.c_double, C.POINTER(VP)]),
    "ckh_mul_no_relin": (C.c_int, [VP, VP, VP, C.POINTER(VP)]),
    "ckh_relinearize": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_mul": (C.c_int, [VP, VP, VP, C.POINTER(VP)]),
    "ckh_rescale": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_mod_switch_to_level": (C.c_int, [VP, VP, u32, C.POINTER(VP)]),
    "ckh_relabel": (C.c_int, [VP, VP, C.c_double, C.POINTER(VP)]),
    "ckh_rotate": (C.c_int, [VP, VP, C.c_int32, C.POINTER(VP)]),
    "ckh_conjugate": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_hoisted_rotate": (C.c_int, [VP, VP, C.POINTER(C.c_int32), u32, C.POINTER(VP)]),
    "ckh_relinearize_batch": (C.c_int, [VP, u32, C.POINTER(VP), C.POINTER(VP)]),
    "ckh_matvec_plain": (C.c_int, [VP, VP, C.POINTER(VP), C.c_double, u32, u32, u32, C.c_int32, C.POINTER(VP)]),
    "ckh_matvec_enc": (C.c_int, [VP, VP, C.POINTER(VP), u32, u32, u32, C.c_int32, C.POINTER(VP)]),
    "ckh_evaluator_set_config": (C.c_int, [VP, u32, u32, u32]),
    "ckh_evaluator_get_config": (C.c_int, [VP, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
    "ckh_mul_scalar": (C.c_int, [VP, VP, C.c_double, C.c_double, C.POINTER(VP)]),
    "ckh_mul_plain_to_scale": (C.c_int, [VP, VP, VP, C.c_double, C.c_double, C.POINTER(VP)]),
    "ckh_exact_product_scale": (C.c_int, [VP, u32, C.c_double, C.POINTER(C.c_double)]),
    "ckh_target_plain_scale": (C.c_int, [VP, VP, C.c_double, C.POINTER(C.c_double)]),
    "ckh_mul_lazy": (C.c_int, [VP, VP, VP, C.POINTER(VP)]),
    "ckh_pool_used_memory": (C.c_int, [VP, C.POINTER(u64)]),
    "ckh_raise_from_level0": (C.c_int, [VP, VP, u32, C.POINTER(VP)]),
    "ckh_mod_raise": (C.c_int, [VP, VP, VP, VP, C.POINTER(VP)]),
    "ckh_mod_raise_main_key": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_matvec_enc_batch": (C.c_int, [VP, u32, C.POINTER(VP), C.POINTER(C.POINTER(VP)), C.c_void_p,
                                       C.POINTER(VP)]),
    "ckh_matvec_enc_workspace_bytes": (C.c_int, [VP, u32, u32, u32, C.POINTER(u64)]),
    "ckh_coeff_to_slot": (C.c_int, [VP, VP, u32, C.POINTER(LinearStageC), C.POINTER(VP), C.POINTER(VP)]),
    "ckh_slot_to_coeff": (C.c_int, [VP, VP, VP, u32, C.POINTER(LinearStageC), C.POINTER(VP)]),
    "ckh_matvec_plain_sum": (C.c_int, [VP, u32, C.POINTER(VP), C.POINTER(C.POINTER(VP)), C.c_double, u32, u32,
                                       u32, C.c_int32, C.POINTER(VP)]),
    "ckh_set_pool_release_threshold": (C.c_int, [VP, u64]),
    "ckh_trim_pool": (C.c_int, [VP, u64]),
    "ckh_relin_lazy": (C.c_int, [VP, VP, C.POINTER(VP)]),
    "ckh_server_create": (C.c_int, [C.POINTER(ServerConfig), C.POINTER(VP)]),
    "ckh_server_destroy": (None, [VP]),
    "ckh_server_now": (C.c_double, [VP]),
    "ckh_server_submit": (C.c_int, [VP, C.POINTER(Request), P64]),
    "ckh_server_poll": (C.c_int, [VP, u64, C.c_int, C.POINTER(C.c_int), C.POINTER(VP), u32, C.POINTER(u32),
                                  C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "ckh_server_stats_get": (C.c_int, [VP, C.POINTER(ServerStats)]),
}

_lib = None


def lib():
    """The loaded library (raises if libckh.so is missing: no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return sorted(_SIGS)


def _check(status):
    if status != 0:
        msg = lib().ckh_last_error().decode(errors="replace")
        raise _ERRORS.get(status, HekitError)(msg)


def _ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P64)


def params_primes(log_n, q_bits, p_bits):
    q = np.zeros(len(q_bits), dtype=np.uint64)
    p = np.zeros(len(p_bits), dtype=np.uint64)
    qb = (u32 * len(q_bits))(*q_bits)
    pb = (u32 * len(p_bits))(*p_bits)
    _check(lib().ckh_params_primes(log_n, qb, len(q_bits), pb, len(p_bits), _ptr(q), _ptr(p)))
    return [int(x) for x in q], [int(x) for x in p]


def galois_elt(log_n, step):
    return lib().ckh_galois_elt(log_n, step)


def conjugation_elt(log_n):
    return lib().ckh_conjugation_elt(log_n)


class Context:
    def __init__(self, log_n, q, p, dnum, ordinal=0, n_streams=1):
        self._q = np.ascontiguousarray(np.asarray(q, dtype=np.uint64))
        self._p = np.ascontiguousarray(np.asarray(p, dtype=np.uint64))
        prm = Params(log_n, _ptr(self._q), len(q), _ptr(self._p), len(p), dnum)
        h = VP()
        _check(lib().ckh_device_open(ordinal, C.byref(prm), n_streams, C.byref(h)))
        self.h = h
        self.log_n, self.n, self.dnum = log_n, 1 << log_n, dnum
        self.q, self.p = [int(x) for x in q], [int(x) for x in p]
        self.alpha = len(q) // dnum

    def close(self):
        if getattr(self, "h", None):
            lib().ckh_device_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def k(self):
        return len(self.p)

    def synchronize(self):
        _check(lib().ckh_synchronize(self.h))

    def autotune(self):
        import json
        buf = C.create_string_buffer(8192)
        _check(lib().ckh_autotune(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def autotune_ks(self, polys, key):
        """Picks the key-switch schedule on this batch (polys under key); returns the timings."""
        import json
        buf = C.create_string_buffer(4096)
        arr = (VP * len(polys))(*[p.h for p in polys])
        _check(lib().ckh_autotune_ks(self.h, len(polys), arr, key.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def pool_used_memory(self):
        v = u64(0)
        _check(lib().ckh_pool_used_memory(self.h, C.byref(v)))
        return v.value

    def set_pool_release_threshold(self, nbytes):
        _check(lib().ckh_set_pool_release_threshold(self.h, nbytes))

    def trim_pool(self, keep_bytes=0):
        _check(lib().ckh_trim_pool(self.h, keep_bytes))

    def set_tuning(self, key, value):
        _check(lib().ckh_set_tuning(self.h, key.encode(), value))

    def tuning(self, key):
        v = C.c_int64(0)
        _check(lib().ckh_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def graph_stats(self):
        """(captured graphs, graph replays) of this context."""
        a, b = u64(0), u64(0)
        _check(lib().ckh_graph_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_stream(self, i):
        _check(lib().ckh_set_stream(self.h, i))

    def stream_handle(self):
        return lib().ckh_stream_handle(self.h)

    def memory_info(self):
        a, b = u64(0), u64(0)
        _check(lib().ckh_memory_info(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # polynomials
    def upload(self, limbs, level, p_limbs=0, ntt=True):
        a = np.ascontiguousarray(np.asarray(limbs, dtype=np.uint64))
        assert a.size == (level + 1 + p_limbs) * self.n
        h = VP()
        _check(lib().ckh_poly_upload(self.h, _ptr(a), level, p_limbs, int(ntt), C.byref(h)))
        return Poly(self, h)

    def uniform(self, level, seed, stream0=0, p_limbs=0, ntt=True):
        h = VP()
        _check(lib().ckh_poly_uniform(self.h, level, p_limbs, int(ntt), seed, stream0, C.byref(h)))
        return Poly(self, h)

    def ksk_upload(self, limbs, max_level, galois=0):
        a = np.ascontiguousarray(np.asarray(limbs, dtype=np.uint64))
        h = VP()
        _check(lib().ckh_ksk_upload(self.h, _ptr(a), max_level, galois, C.byref(h)))
        return Ksk(self, h)

    def ksk_uniform(self, max_level, seed, galois=0):
        h = VP()
        _check(lib().ckh_ksk_uniform(self.h, max_level, galois, seed, C.byref(h)))
        return Ksk(self, h)

    def decompose(self, d):
        arr = (VP * self.dnum)()
        cnt = u32(0)
        _check(lib().ckh_decompose(self.h, d.h, arr, C.byref(cnt)))
        return [Poly(self, VP(arr[i])) for i in range(cnt.value)]

    def apply_key(self, digits, key):
        arr = (VP * len(digits))(*[d.h for d in digits])
        kb, ka = VP(), VP()
        _check(lib().ckh_apply_key(self.h, arr, len(digits), key.h, C.byref(kb), C.byref(ka)))
        return Poly(self, kb), Poly(self, ka)

    def key_switch(self, d, key):
        kb, ka = VP(), VP()
        _check(lib().ckh_key_switch(self.h, d.h, key.h, C.byref(kb), C.byref(ka)))
        return Poly(self, kb), Poly(self, ka)

    def key_switch_batch(self, ds, key):
        n = len(ds)
        arr = (VP * n)(*[d.h for d in ds])
        kb, ka = (VP * n)(), (VP * n)()
        _check(lib().ckh_key_switch_batch(self.h, n, arr, key.h, kb, ka))
        return [(Poly(self, VP(kb[i])), Poly(self, VP(ka[i]))) for i in range(n)]

    def key_switch_batch_static(self, ds, key, outs):
        """The batched switch into existing outputs [(kb, ka), ...] (overwritten in place), with
        no device allocation: intermediates in the GPU's fixed workspace."""
        n = len(ds)
        arr = (VP * n)(*[d.h for d in ds])
        kb = (VP * n)(*[o[0].h for o in outs])
        ka = (VP * n)(*[o[1].h for o in outs])
        _check(lib().ckh_key_switch_batch_static(self.h, n, arr, key.h, kb, ka))

    def time_copy(self, nbytes, reps=5):
        """Mean ms of a streaming device copy of nbytes (16-byte words, HIP events on this
        context's stream)."""
        ms = C.c_double()
        _check(lib().ckh_time_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def time_ntt(self, polys, forward=True, reps=20):
        """Mean ms per batched in-place NTT launch over every limb of `polys` (HIP events on
        this context's stream). Overwrites the polynomials."""
        arr = (VP * len(polys))(*[p.h for p in polys])
        ms = C.c_double()
        _check(lib().ckh_time_ntt(self.h, arr, len(polys), int(forward), reps, C.byref(ms)))
        return ms.value

    def ct_upload(self, polys, level, scale):
        a = np.ascontiguousarray(np.asarray(polys, dtype=np.uint64))
        ncomp = a.size // ((level + 1) * self.n)
        h = VP()
        _check(lib().ckh_ct_upload(self.h, _ptr(a), ncomp, level, scale, C.byref(h)))
        return Ct(self, h)

    def ct_from_polys(self, polys, scale):
        arr = (VP * len(polys))(*[p.h for p in polys])
        h = VP()
        _check(lib().ckh_ct_from_polys(self.h, arr, len(polys), scale, C.byref(h)))
        return Ct(self, h)


class Poly:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ckh_poly_release(self.h)
            self.h = None

    def info(self):
        lv, pl, nt = u32(0), u32(0), C.c_int(0)
        _check(lib().ckh_poly_info(self.h, C.byref(lv), C.byref(pl), C.byref(nt)))
        return lv.value, pl.value, bool(nt.value)

    @property
    def level(self):
        return self.info()[0]

    def download(self):
        lv, pl, _ = self.info()
        out = np.zeros((lv + 1 + pl, self.ctx.n), dtype=np.uint64)
        _check(lib().ckh_poly_download(self.ctx.h, self.h, _ptr(out)))
        return out

    def retain(self):
        return Poly(self.ctx, VP(lib().ckh_poly_retain(self.h)))

    def digest(self):
        d = u64(0)
        _check(lib().ckh_poly_digest(self.ctx.h, self.h, C.byref(d)))
        return d.value

    def ntt(self):
        h = VP()
        _check(lib().ckh_ntt(self.ctx.h, self.h, C.byref(h)))
        return Poly(self.ctx, h)

    def intt(self):
        h = VP()
        _check(lib().ckh_intt(self.ctx.h, self.h, C.byref(h)))
        return Poly(self.ctx, h)

    def automorphism(self, g):
        h = VP()
        _check(lib().ckh_automorphism(self.ctx.h, self.h, g, C.byref(h)))
        return Poly(self.ctx, h)

    def rescale_by_last(self):
        h = VP()
        _check(lib().ckh_rescale_poly(self.ctx.h, self.h, C.byref(h)))
        return Poly(self.ctx, h)

    def mod_down(self):
        h = VP()
        _check(lib().ckh_mod_down(self.ctx.h, self.h, C.byref(h)))
        return Poly(self.ctx, h)

    def raise_from_level0(self, top):
        h = VP()
        _check(lib().ckh_raise_from_level0(self.ctx.h, self.h, top, C.byref(h)))
        return Poly(self.ctx, h)


class Ksk:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ckh_ksk_release(self.h)
            self.h = None

    def info(self):
        lv, g, b = u32(0), u64(0), u64(0)
        _check(lib().ckh_ksk_info(self.h, C.byref(lv), C.byref(g), C.byref(b)))
        return lv.value, g.value, b.value


class Ct:
    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ckh_ct_release(self.h)
            self.h = None

    def info(self):
        nc, lv, sc = u32(0), u32(0), C.c_double(0)
        _check(lib().ckh_ct_info(self.h, C.byref(nc), C.byref(lv), C.byref(sc)))
        return nc.value, lv.value, sc.value

    @property
    def level(self):
        return self.info()[1]

    @property
    def scale(self):
        return self.info()[2]

    def size(self):
        return self.info()[0]

    def download(self):
        nc, lv, _ = self.info()
        out = np.zeros((nc, lv + 1, self.ctx.n), dtype=np.uint64)
        _check(lib().ckh_ct_download(self.ctx.h, self.h, _ptr(out)))
        return out

    def component(self, i):
        return Poly(self.ctx, VP(lib().ckh_ct_component(self.h, i)))

    def digests(self):
        return tuple(self.component(i).digest() for i in range(self.size()))


class Evaluator:
    """HeOps over the device (method names as in crates/ckks-train-ckks/src/ops.rs)."""

    def __init__(self, ctx, relin=None, galois=()):
        self.ctx = ctx
        self._keys = [relin] + list(galois)  # keep the key handles alive
        arr = (VP * max(len(galois), 1))(*[g.h for g in galois])
        h = VP()
        _check(lib().ckh_evaluator_create(ctx.h, relin.h if relin else None, arr, len(galois), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ckh_evaluator_destroy(self.h)
            self.h = None

    def fork(self):
        e = Evaluator.__new__(Evaluator)
        e.ctx, e._keys = self.ctx, self._keys
        h = VP()
        _check(lib().ckh_evaluator_fork(self.h, C.byref(h)))
        e.h = h
        return e

    def counter(self):
        c = OpCounter()
        _check(lib().ckh_counter(self.h, C.byref(c)))
        return c

    def reset_counter(self):
        _check(lib().ckh_reset_counter(self.h))

    def add_counts(self, c):
        _check(lib().ckh_add_counts(self.h, C.byref(c)))

    def has_relin_key(self):
        return bool(lib().ckh_has_relin_key(self.h))

    def has_galois_key(self, step):
        return bool(lib().ckh_has_galois_key(self.h, step))

    def owns(self, ct):
        return bool(lib().ckh_owns(self.h, ct.h))

    def set_config(self, hoist=True, bsgs=True, lazy_relin=True):
        """Evaluator::with_config (EvalConfig, ops.rs:111-134)."""
        _check(lib().ckh_evaluator_set_config(self.h, int(hoist), int(bsgs), int(lazy_relin)))

    def config(self):
        a, b, c = u32(0), u32(0), u32(0)
        _check(lib().ckh_evaluator_get_config(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"hoist": bool(a.value), "bsgs": bool(b.value), "lazy_relin": bool(c.value)}

    def exact_product_scale(self, level, target_scale):
        v = C.c_double(0)
        _check(lib().ckh_exact_product_scale(self.h, level, target_scale, C.byref(v)))
        return v.value

    def target_plain_scale(self, ct, target_scale):
        v = C.c_double(0)
        _check(lib().ckh_target_plain_scale(self.h, ct.h, target_scale, C.byref(v)))
        return v.value

    def mul_scalar(self, ct, value, target_scale):
        return self._op(lib().ckh_mul_scalar, ct.h, value, target_scale)

    def mul_plain_to_scale(self, a, pt, pt_scale, target_scale):
        return self._op(lib().ckh_mul_plain_to_scale, a.h, pt.h, pt_scale, target_scale)

    def mod_raise(self, ct, to_sparse, to_dense):
        return self._op(lib().ckh_mod_raise, ct.h, to_sparse.h, to_dense.h)

    def mod_raise_main_key(self, ct):
        return self._op(lib().ckh_mod_raise_main_key, ct.h)

    def mul_lazy(self, a, b):
        return self._op(lib().ckh_mul_lazy, a.h, b.h)

    def relin_lazy(self, s):
        return self._op(lib().ckh_relin_lazy, s.h)

    def _op(self, fn, *args):
        h = VP()
        _check(fn(self.h, *args, C.byref(h)))
        return Ct(self.ctx, h)

    def add(self, a, b):
        return self._op(lib().ckh_add, a.h, b.h)

    def sub(self, a, b):
        return self._op(lib().ckh_sub, a.h, b.h)

    def negate(self, a):
        return self._op(lib().ckh_negate, a.h)

    def add_plain(self, a, pt, pt_scale):
        return self._op(lib().ckh_add_plain, a.h, pt.h, pt_scale)

    def mul_plain(self, a, pt, pt_scale):
        return self._op(lib().ckh_mul_plain, a.h, pt.h, pt_scale)

    def mul_no_relin(self, a, b):
        return self._op(lib().ckh_mul_no_relin, a.h, b.h)

    def relinearize(self, ct):
        return self._op(lib().ckh_relinearize, ct.h)

    def mul(self, a, b):
        return self._op(lib().ckh_mul, a.h, b.h)

    def rescale(self, ct):
        return self._op(lib().ckh_rescale, ct.h)

    def mod_switch_to_level(self, ct, level):
        return self._op(lib().ckh_mod_switch_to_level, ct.h, level)

    def relabel(self, ct, scale):
        return self._op(lib().ckh_relabel, ct.h, scale)

    def rotate(self, ct, steps):
        return self._op(lib().ckh_rotate, ct.h, steps)

    def conjugate(self, ct):
        return self._op(lib().ckh_conjugate, ct.h)

    def hoisted_rotate(self, ct, steps):
        n = len(steps)
        st = (C.c_int32 * max(n, 1))(*steps)
        outs = (VP * max(n, 1))()
        _check(lib().ckh_hoisted_rotate(self.h, ct.h, st, n, outs))
        return [Ct(self.ctx, VP(outs[i])) for i in range(n)]

    def matvec_plain(self, x, entries, pt_scale, baby, giant, stride=1, offset=0):
        """BSGS product; entries[k1 * baby + k2] is a Poly or None (zero entry), pre-rotated by
        -(offset + baby k1) stride."""
        arr = (VP * len(entries))(*[e.h if e is not None else None for e in entries])
        return self._op(lib().ckh_matvec_plain, x.h, arr, pt_scale, baby, giant, stride, offset)

    def matvec_enc(self, x, entries, baby, giant, stride=1, offset=0):
        """BSGS product with encrypted diagonals (Ct entries or None)."""
        arr = (VP * len(entries))(*[e.h if e is not None else None for e in entries])
        return self._op(lib().ckh_matvec_enc, x.h, arr, baby, giant, stride, offset)

    def matvec_plain_sum(self, xs, entries, pt_scale, baby, giant, stride=1, offset=0):
        """sum_t W_t x_t; entries[t] is a list of Poly or None."""
        n = len(xs)
        xa = (VP * max(n, 1))(*[x.h for x in xs])
        rows = [(VP * len(e))(*[p.h if p is not None else None for p in e]) for e in entries]
        ea = (C.POINTER(VP) * max(n, 1))(*[C.cast(r, C.POINTER(VP)) for r in rows])
        return self._op(lib().ckh_matvec_plain_sum, n, xa, ea, pt_scale, baby, giant, stride, offset)

    @staticmethod
    def _stages(stages):
        """The stages as C structs, and the arrays they point into (kept alive by the caller)."""
        keep = []
        arr = (LinearStageC * max(len(stages), 1))()
        for i, st in enumerate(stages):
            d = (VP * max(len(st.diags), 1))(*[p.h if p is not None else None for p in st.diags])
            keep.append(d)
            sec = None
            if st.second is not None:
                sec = (VP * max(len(st.second), 1))(*[p.h if p is not None else None for p in st.second])
                keep.append(sec)
            arr[i] = LinearStageC(st.direction, st.level, st.scale, BsgsShape(*st.shape), C.cast(d, C.POINTER(VP)),
                                  st.pt_scale, BsgsShape(*st.second_shape),
                                  C.cast(sec, C.POINTER(VP)) if sec is not None else None, st.second_pt_scale)
        return arr, keep

    def coeff_to_slot(self, ct, stages):
        """Evaluator::coeff_to_slot (linear.rs:776-806): (re, im)."""
        arr, keep = self._stages(stages)
        re, im = VP(), VP()
        _check(lib().ckh_coeff_to_slot(self.h, ct.h, len(stages), arr, C.byref(re), C.byref(im)))
        del keep
        return Ct(self.ctx, re), Ct(self.ctx, im)

    def slot_to_coeff(self, re, im, stages):
        """Evaluator::slot_to_coeff (linear.rs:824-859)."""
        arr, keep = self._stages(stages)
        out = self._op(lib().ckh_slot_to_coeff, re.h, im.h, len(stages), arr)
        del keep
        return out

    def matvec_enc_batch(self, xs, entries, shapes):
        """jobs (x, entries, (baby, giant, stride, offset)); results in job order."""
        n = len(xs)
        xa = (VP * max(n, 1))(*[x.h for x in xs])
        rows = [(VP * len(e))(*[c.h if c is not None else None for c in e]) for e in entries]
        ea = (C.POINTER(VP) * max(n, 1))(*[C.cast(r, C.POINTER(VP)) for r in rows])
        sa = (BsgsShape * max(n, 1))(*[BsgsShape(*sh) for sh in shapes])
        outs = (VP * max(n, 1))()
        _check(lib().ckh_matvec_enc_batch(self.h, n, xa, ea, C.cast(sa, C.c_void_p), outs))
        return [Ct(self.ctx, VP(outs[i])) for i in range(n)]

    def matvec_enc_workspace_bytes(self, level, baby, giant):
        v = u64(0)
        _check(lib().ckh_matvec_enc_workspace_bytes(self.h, level, baby, giant, C.byref(v)))
        return v.value

    def relinearize_batch(self, cts):
        n = len(cts)
        arr = (VP * max(n, 1))(*[c.h for c in cts])
        outs = (VP * max(n, 1))()
        _check(lib().ckh_relinearize_batch(self.h, n, arr, outs))
        return [Ct(self.ctx, VP(outs[i])) for i in range(n)]


class Server:
    """The serving loop (include/ckh.h, "serving")."""

    def __init__(self, hbm_budget_bytes, max_batch=64, max_inflight=0):
        cfg = ServerConfig(hbm_budget_bytes, max_batch, max_inflight)
        h = VP()
        _check(lib().ckh_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if getattr(self, "h", None):
            lib().ckh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def now(self):
        return lib().ckh_server_now(self.h)

    def submit(self, kind, ev, a, b=None, steps=(), arrival_s=-1.0, inject_failure=False):
        st = (C.c_int32 * max(len(steps), 1))(*steps)
        r = Request(kind, ev.h, a.h, b.h if b is not None else None, st, len(steps), arrival_s, int(inject_failure))
        t = u64(0)
        _check(lib().ckh_server_submit(self.h, C.byref(r), C.byref(t)))
        self._keep[t.value] = (ev, a, b)
        return t.value

    def poll(self, ticket, wait=True, max_outs=64):
        """None if not finished, else (status, message, outputs, latency_ms)."""
        fin, n, st, lat = C.c_int(0), u32(0), C.c_int(0), C.c_double(0)
        outs = (VP * max_outs)()
        _check(lib().ckh_server_poll(self.h, ticket, int(wait), C.byref(fin), outs, max_outs, C.byref(n),
                                     C.byref(st), C.byref(lat)))
        if not fin.value:
            return None
        ev, a, b = self._keep.pop(ticket)
        msg = lib().ckh_last_error().decode(errors="replace") if st.value else ""
        cts = [Ct(a.ctx, VP(outs[i])) for i in range(min(n.value, max_outs))]
        return st.value, msg, cts, lat.value

    def stats(self):
        s = ServerStats()
        _check(lib().ckh_server_stats_get(self.h, C.byref(s)))
        return s
