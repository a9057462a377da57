"""ctypes binding of the CPU oracle (oracle/ckks_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline
leg of bench.py, always as the checker. The product library never imports it.
"""
import ctypes as C
import math
import os
import struct
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libckks_oracle.so")
_lib = None

u64 = C.c_uint64
P64 = C.POINTER(C.c_uint64)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_params_new.restype = C.c_void_p
        L.or_params_new.argtypes = [C.c_uint, C.POINTER(C.c_uint), u64, C.POINTER(C.c_uint), u64, u64]
        L.or_params_from_primes.restype = C.c_void_p
        L.or_params_from_primes.argtypes = [C.c_uint, P64, u64, P64, u64, u64]
        L.or_params_free.argtypes = [C.c_void_p]
        L.or_params_q.restype = u64
        L.or_params_q.argtypes = [C.c_void_p, u64]
        L.or_params_p.restype = u64
        L.or_params_p.argtypes = [C.c_void_p, u64]
        for name in ("or_mod_add", "or_mod_sub", "or_mod_mul", "or_mod_pow"):
            f = getattr(L, name)
            f.restype = u64
            f.argtypes = [C.c_void_p, u64, u64]
        L.or_mod_new.argtypes = [u64, C.c_void_p]
        L.or_mod_inv.argtypes = [C.c_void_p, u64, P64]
        L.or_shoup.restype = u64
        L.or_shoup.argtypes = [C.c_void_p, u64]
        L.or_mul_shoup.restype = u64
        L.or_mul_shoup.argtypes = [C.c_void_p, u64, u64, u64]
        L.or_mul_shoup_lazy.restype = u64
        L.or_mul_shoup_lazy.argtypes = [C.c_void_p, u64, u64, u64]
        L.or_mod_barrett.argtypes = [u64, P64, P64]
        L.or_is_prime.argtypes = [u64]
        L.or_ntt_primes_skip.argtypes = [C.c_uint, u64, u64, u64, P64]
        L.or_ntt_forward_limbs.argtypes = [u64, P64, u64, P64]
        L.or_ntt_inverse_limbs.argtypes = [u64, P64, u64, P64]
        L.or_ntt_psi.restype = u64
        L.or_ntt_psi.argtypes = [u64, u64]
        L.or_galois_elt.restype = u64
        L.or_galois_elt.argtypes = [u64, C.c_int]
        L.or_conjugation_elt.restype = u64
        L.or_conjugation_elt.argtypes = [u64]
        L.or_galois_perm_ntt.argtypes = [u64, u64, C.POINTER(C.c_uint32)]
        L.or_automorphism_ntt.argtypes = [u64, u64, P64, P64, u64]
        L.or_automorphism_coeff.argtypes = [u64, u64, P64, P64, P64, u64]
        L.or_convert.argtypes = [P64, u64, P64, u64, u64, P64, P64, C.c_int]
        L.or_rescale_by_last.argtypes = [C.c_void_p, P64, u64, C.c_int]
        L.or_mod_down.argtypes = [C.c_void_p, P64, u64, P64, u64, C.c_int]
        L.or_digit_ranges.restype = u64
        L.or_digit_ranges.argtypes = [C.c_void_p, u64, P64, P64]
        L.or_decompose.argtypes = [C.c_void_p, P64, u64, P64]
        L.or_apply_key.argtypes = [C.c_void_p, P64, u64, u64, P64, u64, P64, P64]
        L.or_key_switch.argtypes = [C.c_void_p, P64, u64, P64, u64, P64, P64]
        L.or_relinearize.argtypes = [C.c_void_p, P64, P64, P64, u64, P64, u64, P64, P64]
        L.or_apply_galois.argtypes = [C.c_void_p, P64, P64, u64, u64, P64, u64, P64, P64]
        L.or_hoisted_galois.argtypes = [C.c_void_p, P64, P64, u64, u64, P64, C.POINTER(P64), u64, P64]
        L.or_rescale_ct.argtypes = [C.c_void_p, P64, u64, u64]
        L.or_mul_no_relin.argtypes = [C.c_void_p, P64, P64, P64, P64, u64, P64, P64, P64]
        L.or_poly_add.argtypes = [C.c_void_p, P64, P64, u64]
        L.or_poly_mul.argtypes = [C.c_void_p, P64, P64, u64]
        L.or_poly_mul_scalar.argtypes = [C.c_void_p, P64, u64, u64, C.c_int]
        L.or_raise_from_level0.argtypes = [C.c_void_p, P64, u64, C.c_int, P64]
        L.or_params_ntt.argtypes = [C.c_void_p, P64, u64, C.c_int]
        L.or_uniform_word.restype = u64
        L.or_uniform_word.argtypes = [u64, u64, u64, u64]
        L.or_fill_uniform.argtypes = [P64, u64, u64, u64, u64]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_get_threads.restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(P64)


def arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


class Mod:
    """A Modulus (q < 2^60) with the oracle's Barrett / Shoup routines."""

    def __init__(self, q):
        self.buf = (u64 * 3)()
        if lib().or_mod_new(q, C.byref(self.buf)) != 0:
            raise ValueError(f"bad modulus {q}")
        self.q = q

    def mul(self, a, b):
        return lib().or_mod_mul(C.byref(self.buf), a, b)

    def inv(self, a):
        out = u64(0)
        ok = lib().or_mod_inv(C.byref(self.buf), a, C.byref(out))
        return out.value if ok else None

    def shoup(self, w):
        return lib().or_shoup(C.byref(self.buf), w)

    def mul_shoup(self, a, w):
        return lib().or_mul_shoup(C.byref(self.buf), a, w % self.q, self.shoup(w))

    def mul_shoup_lazy(self, a, w):
        return lib().or_mul_shoup_lazy(C.byref(self.buf), a, w % self.q, self.shoup(w))


def barrett(q):
    lo, hi = u64(0), u64(0)
    lib().or_mod_barrett(q, C.byref(lo), C.byref(hi))
    return lo.value | (hi.value << 64)


def is_prime(n):
    return bool(lib().or_is_prime(n))


def ntt_primes(bits, n, count, skip=0):
    out = np.zeros(max(count, 1), dtype=np.uint64)
    if lib().or_ntt_primes_skip(bits, n, skip, count, ptr(out)) != 0:
        raise ValueError("InvalidParams")
    return [int(x) for x in out[:count]]


def ntt_forward(n, primes, a):
    a = arr(a).copy()
    q = arr(primes)
    assert lib().or_ntt_forward_limbs(n, ptr(q), len(primes), ptr(a)) == 0
    return a


def ntt_inverse(n, primes, a):
    a = arr(a).copy()
    q = arr(primes)
    assert lib().or_ntt_inverse_limbs(n, ptr(q), len(primes), ptr(a)) == 0
    return a


def psi(n, q):
    return lib().or_ntt_psi(n, q)


def galois_elt(n, step):
    return lib().or_galois_elt(n, step)


def conjugation_elt(n):
    return lib().or_conjugation_elt(n)


def galois_perm(n, g):
    out = np.zeros(n, dtype=np.uint32)
    lib().or_galois_perm_ntt(n, g, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def automorphism_ntt(n, g, a):
    a = arr(a)
    out = np.zeros_like(a)
    lib().or_automorphism_ntt(n, g, ptr(a), ptr(out), a.size // n)
    return out


def automorphism_coeff(n, g, primes, a):
    a = arr(a)
    out = np.zeros_like(a)
    q = arr(primes)
    lib().or_automorphism_coeff(n, g, ptr(q), ptr(a), ptr(out), a.size // n)
    return out


def convert(src_primes, dst_primes, src, exact):
    src = arr(src)
    k, t = len(src_primes), len(dst_primes)
    n = src.size // k
    out = np.zeros(t * n, dtype=np.uint64)
    lib().or_convert(ptr(arr(src_primes)), k, ptr(arr(dst_primes)), t, n, ptr(src), ptr(out), int(exact))
    return out.reshape(t, n)


def uniform(n, q, seed, stream):
    out = np.zeros(n, dtype=np.uint64)
    lib().or_fill_uniform(ptr(out), n, q, seed, stream)
    return out


def uniform_poly(primes, n, seed, stream0=0):
    """Limb i uniform mod primes[i] from stream stream0 + i."""
    return np.stack([uniform(n, q, seed, stream0 + i) for i, q in enumerate(primes)])


def set_threads(n):
    lib().or_set_threads(n)


def get_threads():
    return lib().or_get_threads()


MIN_PLAIN_SCALE = 1048576.0  # evaluator.rs:108
MAX_SCALAR_INT = 4611686018427387904.0  # evaluator.rs:103, 2^62
LABEL_SEARCH_ULPS = 8  # evaluator.rs:100


def _bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _from_bits(b):
    return struct.unpack("<d", struct.pack("<Q", b & ((1 << 64) - 1)))[0]


def exact_product_scale(q, level, target_scale):
    """exact_product_scale (crates/ckks-train-ckks/src/evaluator.rs:738-772): the f64 label L
    with L / q_level == target_scale, searched from target_scale * q_level outwards (candidates
    bits + k, bits - k for k = 0..=8). Raises ValueError(kind) with the HekitError variant."""
    if level > len(q) - 1:
        raise ValueError("InvalidParams")
    if not (math.isfinite(target_scale) and target_scale > 0.0):
        raise ValueError("InvalidParams")
    if level == 0:
        raise ValueError("NoLevelsLeft")
    ql = float(q[level])
    l0 = target_scale * ql
    if not math.isfinite(l0):
        raise ValueError("InvalidParams")
    b = _bits(l0)
    for k in range(LABEL_SEARCH_ULPS + 1):
        for cand in (b + k, b - k):
            lab = _from_bits(cand)
            if lab / ql == target_scale:
                return lab
    raise ValueError("ScaleMismatch")


def scalar_product_parts(q, level, source_scale, value, target_scale):
    """scalar_product_parts (evaluator.rs:775-797): (label, plaintext scale, integer k)."""
    if not math.isfinite(value):
        raise ValueError("InvalidParams")
    label = exact_product_scale(q, level, target_scale)
    if label / source_scale < MIN_PLAIN_SCALE:
        raise ValueError("InvalidParams")
    x = value * (label / source_scale)
    a = abs(x)  # f64::round: half away from zero (a - floor(a) is exact)
    r = float(math.floor(a))
    if a - r >= 0.5:
        r += 1.0
    k = r if x >= 0 else -r
    if abs(k) >= MAX_SCALAR_INT:
        raise ValueError("InvalidParams")
    return label, label / source_scale, k


class Params:
    """or_params: a CkksParams with its Q and P bases."""

    def __init__(self, log_n, q_bits=None, p_bits=None, dnum=1, q=None, p=None):
        L = lib()
        if q is not None:
            self.h = L.or_params_from_primes(log_n, ptr(arr(q)), len(q), ptr(arr(p)), len(p), dnum)
            nq, np_ = len(q), len(p)
        else:
            qb = (C.c_uint * len(q_bits))(*q_bits)
            pb = (C.c_uint * len(p_bits))(*p_bits)
            self.h = L.or_params_new(log_n, qb, len(q_bits), pb, len(p_bits), dnum)
            nq, np_ = len(q_bits), len(p_bits)
        if not self.h:
            raise ValueError("InvalidParams")
        self.log_n, self.n, self.dnum = log_n, 1 << log_n, dnum
        self.q = [L.or_params_q(self.h, i) for i in range(nq)]
        self.p = [L.or_params_p(self.h, i) for i in range(np_)]
        self.alpha = nq // dnum

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_params_free(self.h)
            self.h = None

    @property
    def k(self):
        return len(self.p)

    def digit_ranges(self, level):
        s = np.zeros(64, dtype=np.uint64)
        e = np.zeros(64, dtype=np.uint64)
        cnt = lib().or_digit_ranges(self.h, level, ptr(s), ptr(e))
        return [(int(s[i]), int(e[i])) for i in range(cnt)]

    def rescale_by_last(self, poly, level, ntt=True):
        a = arr(poly).copy()
        lib().or_rescale_by_last(self.h, ptr(a), level, int(ntt))
        return a.reshape(level + 1, self.n)[:level].copy()

    def mod_down(self, qpart, level, ppart, ntt=True):
        a = arr(qpart).copy()
        pp = arr(ppart)
        lib().or_mod_down(self.h, ptr(a), level, ptr(pp), pp.size // self.n, int(ntt))
        return a.reshape(level + 1, self.n)

    def decompose(self, poly, level):
        nd = len(self.digit_ranges(level))
        out = np.zeros((nd, level + 1 + self.k, self.n), dtype=np.uint64)
        lib().or_decompose(self.h, ptr(arr(poly)), level, ptr(out))
        return out

    def apply_key(self, digits, level, key, key_max_level):
        digits = arr(digits)
        nd = digits.size // ((level + 1 + self.k) * self.n)
        kb = np.zeros((level + 1, self.n), dtype=np.uint64)
        ka = np.zeros_like(kb)
        lib().or_apply_key(self.h, ptr(digits), nd, level, ptr(arr(key)), key_max_level, ptr(kb), ptr(ka))
        return kb, ka

    def key_switch(self, poly, level, key, key_max_level):
        kb = np.zeros((level + 1, self.n), dtype=np.uint64)
        ka = np.zeros_like(kb)
        lib().or_key_switch(self.h, ptr(arr(poly)), level, ptr(arr(key)), key_max_level, ptr(kb), ptr(ka))
        return kb, ka

    def relinearize(self, c0, c1, c2, level, key, key_max_level):
        o0 = np.zeros((level + 1, self.n), dtype=np.uint64)
        o1 = np.zeros_like(o0)
        lib().or_relinearize(self.h, ptr(arr(c0)), ptr(arr(c1)), ptr(arr(c2)), level, ptr(arr(key)),
                             key_max_level, ptr(o0), ptr(o1))
        return o0, o1

    def apply_galois(self, c0, c1, level, g, key, key_max_level):
        o0 = np.zeros((level + 1, self.n), dtype=np.uint64)
        o1 = np.zeros_like(o0)
        lib().or_apply_galois(self.h, ptr(arr(c0)), ptr(arr(c1)), level, g, ptr(arr(key)), key_max_level,
                              ptr(o0), ptr(o1))
        return o0, o1

    def hoisted_galois(self, c0, c1, level, elts, keys, key_max_level):
        keys = [arr(k) for k in keys]
        kp = (P64 * len(keys))(*[ptr(k) for k in keys])
        outs = np.zeros((len(elts), 2, level + 1, self.n), dtype=np.uint64)
        lib().or_hoisted_galois(self.h, ptr(arr(c0)), ptr(arr(c1)), level, len(elts), ptr(arr(elts)), kp,
                                key_max_level, ptr(outs))
        return outs

    def rescale_ct(self, ct, level):
        a = arr(ct).copy()
        ncomp = a.size // ((level + 1) * self.n)
        lib().or_rescale_ct(self.h, ptr(a), ncomp, level)
        return a.reshape(ncomp, level + 1, self.n)[:, :level].copy()

    def mul_no_relin(self, a0, a1, b0, b1, level):
        d = [np.zeros((level + 1, self.n), dtype=np.uint64) for _ in range(3)]
        lib().or_mul_no_relin(self.h, ptr(arr(a0)), ptr(arr(a1)), ptr(arr(b0)), ptr(arr(b1)), level,
                              ptr(d[0]), ptr(d[1]), ptr(d[2]))
        return d

    def poly_add(self, a, b, level):
        out = arr(a).copy()
        lib().or_poly_add(self.h, ptr(out), ptr(arr(b)), level + 1)
        return out.reshape(level + 1, self.n)

    def ntt(self, a, limbs, forward=True):
        """Q limbs 0..limbs transformed with the parameters' tables (a copy)."""
        out = arr(a).copy()
        lib().or_params_ntt(self.h, ptr(out), limbs, int(forward))
        return out.reshape(limbs, self.n)

    def raise_from_level0(self, a, top, ntt=True):
        out = np.zeros((top + 1, self.n), dtype=np.uint64)
        lib().or_raise_from_level0(self.h, ptr(arr(a)), top, int(ntt), ptr(out))
        return out

    def poly_mul_scalar(self, a, level, k, neg):
        out = arr(a).copy()
        lib().or_poly_mul_scalar(self.h, ptr(out), level + 1, k, int(neg))
        return out.reshape(level + 1, self.n)

    def exact_product_scale(self, level, target_scale):
        return exact_product_scale(self.q, level, target_scale)

    def mul_scalar_parts(self, level, source_scale, value, target_scale):
        return scalar_product_parts(self.q, level, source_scale, value, target_scale)

    def poly_mul(self, a, b, level):
        out = arr(a).copy()
        lib().or_poly_mul(self.h, ptr(out), ptr(arr(b)), level + 1)
        return out.reshape(level + 1, self.n)

    def poly_sub(self, a, b, level):
        """(a - b) mod q_i per limb (Q limbs 0..=level)."""
        a = arr(a).reshape(level + 1, self.n)
        b = arr(b).reshape(level + 1, self.n)
        q = np.array(self.q[: level + 1], dtype=np.uint64)[:, None]
        return np.where(a >= b, a - b, a + (q - b))

    def mul_by_x_half(self, a, level):
        """X^(N/2) a(X) mod (X^N + 1) for an NTT-form polynomial over Q limbs 0..=level
        (Evaluator::mul_by_x_half, crates/ckks-train-ckks/src/bootstrap/linear.rs:921-945, which
        multiplies by the NTT of the monomial), restated from the ring's definition: back to
        coefficients, coefficient i moves to i + N/2 and the upper half wraps around negated,
        then forward again."""
        c = self.ntt(a, level + 1, forward=False)
        h = self.n // 2
        q = np.array(self.q[: level + 1], dtype=np.uint64)[:, None]
        out = np.empty_like(c)
        out[:, h:] = c[:, :h]
        up = c[:, h:]
        out[:, :h] = np.where(up == 0, up, q - up)
        return self.ntt(out, level + 1, forward=True)

    def key_shape(self, key_max_level):
        """(parts, limbs per key polynomial) of a key serving levels 0..=key_max_level."""
        return len(self.digit_ranges(key_max_level)), key_max_level + 1 + self.k

    def uniform_key(self, key_max_level, seed):
        """A key in the device layout with uniform residues (keyswitch arithmetic is
        key-agnostic): per part, b_j then a_j, each Q limbs 0..=key_max_level then P limbs."""
        parts, lp = self.key_shape(key_max_level)
        primes = list(self.q[: key_max_level + 1]) + list(self.p)
        polys = []
        for j in range(2 * parts):
            polys.append(uniform_poly(primes, self.n, seed, stream0=1000 * (j + 1)))
        return np.stack(polys)
