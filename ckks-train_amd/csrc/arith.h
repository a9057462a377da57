// Portable part of the library's shared definitions: limb descriptors, per-prime constants and
// the modular arithmetic of the kernels, written so that a host program can include it too (the
// CPU replay of the kernels in tests/emu). The includer defines CKH_DEV (the device prelude,
// common.h: __device__ __forceinline__), mulhi(a, b) (the high word of a 64 x 64 product) and
// the f64 bit casts bits_f64 / f64_bits before including this header.
//
// Residues are u64 in HBM, limb-major: limb i of a polynomial is N contiguous words. All
// arithmetic is exact modular integer arithmetic, so results are unique residues whatever the
// evaluation order; the one floating-point step (the overflow estimate of the exact base
// conversion) lives in conv_arith.h, and everything is compiled without FMA contraction.
#pragma once

#include <stdint.h>

// Checks of the f64 arithmetic's preconditions and results (empty in the library; the CPU
// replay defines them to verify every product exactly against integer arithmetic).
#ifndef CKH_CHECK_FMULMOD
#define CKH_CHECK_FMULMOD(y, w, wq, q, t) ((void)0)
#endif
#ifndef CKH_CHECK_FREDUCE
#define CKH_CHECK_FREDUCE(x, q, r) ((void)0)
#endif
#ifndef CKH_CHECK_SHOUP
#define CKH_CHECK_SHOUP(a, w, ws, q, r) ((void)0)
#endif

typedef uint64_t u64;
typedef uint32_t u32;

namespace ckh {

// One limb of work: where to read, where to write, and which prime (index into the context's
// prime list: Q primes first, then P primes).
struct LimbRef {
    const u64* src;
    u64* dst;
    u32 prime;
    u32 aux;
};

// A twiddle and its Shoup companion, loaded as one 16-byte word.
struct Tw {
    u64 w, ws;
};

// A twiddle for the f64 path (primes below 2^50): w exactly, and w / q rounded.
struct FTw {
    double w, wq;
};

// out <- (a - t) * inv mod q (+ add when add != null), one limb: finishes rescale and mod_down.
struct FinishJob {
    const u64* a;
    const u64* t;
    u64* out;
    const u64* add;
    u64 inv, inv_s;
    u32 prime, pad;
    FTw inv_f;  // inv and its quotient by q (f64 path, q < 2^50)
};

// One limb of a key-switch inner product fused with the forward row pass (primes below 2^50):
// digit j's limb is the column-pass output at mid + j * mid_stride, except digit own_j whose
// limb is the NTT-form own (not transformed); key part j's b limb at key + 2 j key_poly, its a
// limb key_poly further; acc_b / acc_a receive the sums.
struct RowIpJob {
    const u64* mid;
    u64 mid_stride;
    const u64* own;
    const u64* key;
    u64 key_poly;
    u64* acc_b;
    u64* acc_a;
    u32 ndig, prime, own_j, pad;
};

// The fused row pass + inner product of one limb followed by mod_down's finish of both sums:
// the inner-product sums stay in registers, the limb's converted P part (zb / za, column pass
// done) is row-transformed, and out = (sum - NTT(z)) * P^-1 (+ add) is written
// (keyswitch.rs:254-309, baseconv.rs:459-511); ip.acc_b / ip.acc_a are not used.
struct RowIpFinJob {
    RowIpJob ip;
    const u64* zb;
    const u64* za;
    u64* out_b;
    u64* out_a;
    const u64* add_b;  // optional addends (nullptr: none)
    const u64* add_a;
    FTw inv;  // P^-1 mod q_i and its quotient by q_i (q_i < 2^50)
};

// Row factors per row of the register-radix row pass (one 128-byte line of (w, w/q) pairs).
constexpr u32 ROW_FACTOR_SLOTS = 8;

// Primes below this bound use the f64 arithmetic (every intermediate stays an exact integer
// below 2^53); larger primes use 64-bit integer Shoup arithmetic.
constexpr u64 FP_PRIME_BOUND = 1ull << 50;

// Per-prime constants (device).
struct PrimeConst {
    u64 q;
    u64 n_inv, n_inv_s;     // N^-1 and its Shoup companion
    u64 w_n_inv, w_n_inv_s; // psi^-1 * N^-1 (fused last inverse stage)
    u64 one_s;              // floor(2^64 / q): Shoup companion of 1
    u64 two64, two64_s;     // 2^64 mod q and companion
    double qf, qinv;        // q and 1/q as f64 (f64 path)
    FTw n_inv_f, w_n_inv_f; // the last inverse stage constants, f64 path
};

// The last stage of an inverse NTT whose output limb only feeds a base conversion, folded with
// the conversion's first step y = [x + h]_q (Q / q)^-1 mod q (x the transform's output, h = q >> 1,
// Q the conversion's source modulus): n^-1 (Q / q)^-1 and psi^-1 n^-1 (Q / q)^-1 mod q replace
// the stage's constants and h (Q / q)^-1 mod q is added, so the limb holds y ("y form").
struct YForm {
    u64 n, n_s, wn, wn_s;  // integer path: the two constants and their Shoup companions
    u64 h;                 // h (Q / q)^-1 mod q
    FTw n_f, wn_f;         // f64 path: the constants and their quotients by q
    double h_f;
};

// Shoup product in [0, 2q) for any u64 a (w < q fixed, ws = floor(w 2^64 / q)).
CKH_DEV u64 shoup_lazy(u64 a, u64 w, u64 ws, u64 q) {
    const u64 r = a * w - mulhi(a, ws) * q;
    CKH_CHECK_SHOUP(a, w, ws, q, r);
    return r;
}

CKH_DEV u64 csub(u64 a, u64 q) { return a >= q ? a - q : a; }

CKH_DEV u64 shoup(u64 a, u64 w, u64 ws, u64 q) { return csub(shoup_lazy(a, w, ws, q), q); }

CKH_DEV u64 add_mod(u64 a, u64 b, u64 q) { return csub(a + b, q); }

CKH_DEV u64 sub_mod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

// Any u64 reduced mod q.
CKH_DEV u64 reduce64(u64 a, const PrimeConst& c) { return csub(a - mulhi(a, c.one_s) * c.q, c.q); }

// (hi 2^64 + lo) mod q for any hi, lo.
CKH_DEV u64 reduce128(u64 hi, u64 lo, const PrimeConst& c) {
    return add_mod(shoup(hi, c.two64, c.two64_s, c.q), reduce64(lo, c), c.q);
}

// a * b mod q for reduced a, b.
CKH_DEV u64 mul_mod(u64 a, u64 b, const PrimeConst& c) { return reduce128(mulhi(a, b), a * b, c); }

// ---- f64 arithmetic for primes below 2^50 (FP_PRIME_BOUND): residues as integer-valued f64.

// u64 < 2^53 -> f64, exact.
CKH_DEV double u2d(u64 v) { return __builtin_fma((double)(u32)(v >> 32), 4294967296.0, (double)(u32)v); }

// Non-negative integer-valued f64 below 2^53 -> u64, exact.
CKH_DEV u64 d2u(double d) {
    const u32 hi = (u32)(d * (1.0 / 4294967296.0));
    const double lo = __builtin_fma(-(double)hi, 4294967296.0, d);
    return ((u64)hi << 32) | (u32)lo;
}

// y w mod q as a signed integer-valued f64, for integer-valued y, w with |w| < q < 2^50 and
// wq = w / q rounded once or twice: h = RN(y w) and l = fma(y, w, -h) split the product
// exactly; k = rint(y wq) misses y w / q by at most 1/2 + 3 |y| / (16 q) (the rounding errors
// of wq and of y wq), so t = y w - k q has |t| <= (1/2 + 3 |y| / (16 q)) q, e.g. 0.69 q for
// |y| <= q; while that stays far below 2^53, fma(-k, q, h) + l computes t exactly.
CKH_DEV double fmulmod(double y, double w, double wq, double q) {
    const double h = y * w;
    const double l = __builtin_fma(y, w, -h);
    const double k = __builtin_rint(y * wq);
    const double t = __builtin_fma(-k, q, h) + l;
    CKH_CHECK_FMULMOD(y, w, wq, q, t);
    return t;
}

// x - rint(x / q) q: |result| <= q / 2 (plus rounding slack) for |x| < 2^53.
CKH_DEV double freduce(double x, double q, double qinv) {
    const double r = __builtin_fma(-__builtin_rint(x * qinv), q, x);
    CKH_CHECK_FREDUCE(x, q, r);
    return r;
}

// The canonical residue in [0, q) as f64.
CKH_DEV double fcanon(double x, double q, double qinv) {
    const double r = freduce(x, q, qinv);
    return r < 0.0 ? r + q : r;
}

// Exact sum of products of values below 2^60 (every residue here), from 32-bit halves:
// y c = y0 c0 + (y0 c1 + y1 c0) 2^32 + y1 c1 2^64 with y1, c1 < 2^28, one 32 x 32 -> 64 multiply
// per partial product instead of full 64 x 64 products; the low and middle sums carry into
// counters, the high one (below 2^56 per term) cannot overflow for fewer than 256 terms.
struct Dot60 {
    u64 l = 0, m = 0, h = 0;
    u32 cl = 0, cm = 0;
    CKH_DEV void add(u64 y, u64 c) {
        const u32 y0 = (u32)y, y1 = (u32)(y >> 32), c0 = (u32)c, c1 = (u32)(c >> 32);
        const u64 p = (u64)y0 * c0;
        l += p;
        cl += l < p ? 1u : 0u;
        const u64 t = (u64)y0 * c1 + (u64)y1 * c0;
        m += t;
        cm += m < t ? 1u : 0u;
        h += (u64)y1 * c1;
    }
    // The sum as hi 2^64 + lo.
    CKH_DEV void get(u64& hi, u64& lo) const {
        lo = l + (m << 32);
        hi = h + cl + (m >> 32) + ((u64)cm << 32) + (lo < l ? 1 : 0);
    }
};

}  // namespace ckh
