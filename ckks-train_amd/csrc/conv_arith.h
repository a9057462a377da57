// Centered exact fast base conversion: the tables of a conversion and the per-coefficient
// arithmetic of the vector kernel, written so that a host program can include it too (the CPU
// replay of the conversion in tests/emu). The includer has included arith.h and defines cst(p)
// (a pointer to read-only tables), pc_int(p) (a prime's integer-reduction constants) and glb(p)
// (a global-memory view of a pointer).
//
// For sources q_0..q_{k-1} (product Q) and coefficient-form residues x_j, every target r gets
//   y_j = [x_j + (q_j >> 1)]_{q_j} * (Q/q_j)^-1 mod q_j
//   e   = floor(sum_j f64(y_j) * f64(1/q_j))             (k > 1 only)
//   out = sum_j y_j (Q/q_j mod r) - e (Q mod r) - ((Q-1)/2 mod r)   mod r
// which is BaseConverter::convert_exact([x + h]_Q) - h
// (crates/ckks-train-core/src/baseconv.rs:171-233; keyswitch.rs:140-210; baseconv.rs:258-291,
// 459-511). The modular sums give unique residues in any order. The f64 estimate is the one
// non-modular step: it uses the CPU's exact sequence (round-to-nearest u64->f64 conversion,
// round-to-nearest product with the host-computed reciprocal, round-to-nearest sum in source
// order from +0.0, then floor), so every includer compiles it without FMA contraction.
#pragma once

#include "arith.h"

namespace ckh {

constexpr u32 MAX_SRC = 16;

// One source prime q_j of a conversion from Q = prod q_j.
struct ConvSrc {
    u64 q;
    u64 h;               // q_j >> 1: [x + (Q-1)/2]_Q has residue x + h mod q_j
    u64 qhinv, qhinv_s;  // (Q / q_j)^-1 mod q_j
    double inv;          // 1.0 / f64(q_j), round to nearest (host computed)
    double qf;           // q_j as f64 (f64 path)
    double qhinv_f, qhinv_q;  // (Q / q_j)^-1 and its quotient by q_j, f64 path
    u32 fp, pad;              // q_j < 2^50
};

// One target prime r of a conversion.
struct ConvTgt {
    u32 prime;  // index into the context's prime constants
    u32 limb;   // output limb index within the job's output polynomial
    u64 qmod, qmod_s;  // Q mod r
    u64 hmod;          // (Q - 1)/2 mod r
    u64 qhat[MAX_SRC];                   // (Q / q_j) mod r
    double qhat_f[MAX_SRC], qhat_q[MAX_SRC];      // f64 path: (Q / q_j) mod r and its quotient by r
    double qhat30_f[MAX_SRC], qhat30_q[MAX_SRC];  // split path: 2^30 (Q / q_j) mod r, quotient by r
    double half[2][MAX_SRC];  // piece path: (Q / q_j) mod r in 25-bit halves (low, high)
    FTw p20, p40, p65;        // piece path: 2^20, 2^40, 2^65 mod r (and their quotients by r)
    double qmod_f, qmod_q, hmod_f, rf, rinv;
    // TGT_INT: r at least 2^50, 128-bit integer sums; TGT_FP: r and every source below 2^50, f64
    // products of the sources; TGT_SPLIT: r below 2^50 with larger sources, each source split
    // into 30-bit halves (or, variant 1, 20-bit pieces summed exactly).
    u32 mode, pad;
};

enum : u32 { TGT_INT = 0, TGT_FP = 1, TGT_SPLIT = 2 };

// A conversion: sources [src_first, src_first + nsrc) of the ConvSrc table (their input limbs
// are src_limb0 + j of the job's input), targets [tgt_first, tgt_first + ntgt) of the ConvTgt
// table. exact = use the f64 overflow estimate (nsrc > 1).
struct ConvPlan {
    u32 src_first, nsrc, tgt_first, ntgt, src_limb0, exact;
};

// One launch row: plan `plan`, input polynomial (its limb 0) at `in`, output at `out`.
struct ConvJob {
    const u64* in;
    u64* out;
    u32 plan, pad;
};

// Round-to-nearest u64 -> f64: both halves convert exactly and the sum rounds once.
CKH_DEV double u64_to_f64_rn(u64 v) {
    const double hi = (double)(u32)(v >> 32) * 4294967296.0;
    const double lo = (double)(u32)v;
    return hi + lo;
}

// Coefficient c of one conversion (plan P, input polynomial in_p, output out_p, N = n): the
// output words of targets [t0, t1). MAXS bounds the plan's sources; PIECES selects the exact
// f64 piece sums for 50-bit targets of 60-bit sources. Per-target constants are read through
// cst (the scalar cache on the device); 60-bit targets sum exact 128-bit integer products and
// reduce once.
template <u32 MAXS, bool PIECES>
CKH_DEV void conv_coeff(const ConvSrc* srcs, const ConvTgt* tgts, const PrimeConst* pcs, const ConvPlan& P,
                        const u64* in_p, u64* out_p, u32 n, u32 c, u32 t0, u32 t1) {
    const auto inp = glb(in_p);
    u64 y[MAXS];
    // f64 source: yd = y, yh = 0; integer source: yd / yh = its low / high 30 bits, or (PIECES)
    // yd / yh / yx = its 20-bit pieces from the low end
    double yd[MAXS], yh[MAXS], yx[MAXS];
    double f = 0.0;
    // the piece path needs every source integer (at most 7: the bracketed sums below 2^53)
    bool srcs_int = PIECES && MAXS <= 7;
    for (u32 j = 0; j < P.nsrc && srcs_int; ++j) srcs_int = !cst(srcs + P.src_first + j)->fp;
#pragma unroll
    for (u32 j = 0; j < MAXS; ++j) {
        y[j] = 0;
        yd[j] = 0.0;
        yh[j] = 0.0;
        yx[j] = 0.0;
        if (j < P.nsrc) {
            const ConvSrc* s = srcs + P.src_first + j;
            const u64 x = inp[(u64)(P.src_limb0 + j) * n + c];
            if (s->fp) {
                double sv = u2d(x) + (double)s->h;
                sv = sv >= s->qf ? sv - s->qf : sv;
                double v = fmulmod(sv, s->qhinv_f, s->qhinv_q, s->qf);
                v = v < 0.0 ? v + s->qf : v;
                yd[j] = v;
                y[j] = d2u(v);
            } else {
                y[j] = shoup(add_mod(x, s->h, s->q), s->qhinv, s->qhinv_s, s->q);
                if (srcs_int) {
                    yd[j] = (double)(u32)(y[j] & ((1u << 20) - 1));
                    yh[j] = (double)(u32)((y[j] >> 20) & ((1u << 20) - 1));
                    yx[j] = (double)(u32)(y[j] >> 40);
                } else {
                    yd[j] = (double)(u32)(y[j] & ((1u << 30) - 1));  // low 30 bits, exact
                    yh[j] = (double)(u32)(y[j] >> 30);                // high 30 bits, exact
                }
            }
            if (P.exact) {
                const double prod = u64_to_f64_rn(y[j]) * s->inv;
                f = f + prod;
            }
        }
    }
    const u64 e = P.exact ? (u64)f : 0;  // f >= 0: truncation is floor
    const auto outp = glb(out_p);
    for (u32 t = t0; t < t1; ++t) {
        const auto* g = cst(tgts + P.tgt_first + t);
        u64 o;
        if (PIECES && g->mode == TGT_SPLIT && srcs_int) {
            // y_j = y0 + y1 2^20 + y2 2^40 (20-bit pieces), c_j = c0 + c1 2^25 (25-bit halves):
            // S_ab = sum_j y_ja c_jb is exact (every product < 2^45, at most 7 terms), and
            // sum_j y_j c_j = S00 + (32 S01 + S10) 2^20 + (32 S11 + S20) 2^40 + S21 2^65, the two
            // bracketed sums staying below 2^53.
            const double r = g->rf, rinv = g->rinv;
            double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0, s20 = 0.0, s21 = 0.0;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                const double c0 = g->half[0][j], c1 = g->half[1][j];
                s00 = __builtin_fma(yd[j], c0, s00);
                s01 = __builtin_fma(yd[j], c1, s01);
                s10 = __builtin_fma(yh[j], c0, s10);
                s11 = __builtin_fma(yh[j], c1, s11);
                s20 = __builtin_fma(yx[j], c0, s20);
                s21 = __builtin_fma(yx[j], c1, s21);
            }
            const double t1 = __builtin_fma(s01, 32.0, s10), t2 = __builtin_fma(s11, 32.0, s20);
            double acc = s00 + fmulmod(t1, g->p20.w, g->p20.wq, r);
            acc = acc + fmulmod(t2, g->p40.w, g->p40.wq, r);
            acc = acc + fmulmod(s21, g->p65.w, g->p65.wq, r);
            if (P.exact) acc = acc - fmulmod((double)e, g->qmod_f, g->qmod_q, r);
            acc = acc - g->hmod_f;
            o = d2u(fcanon(acc, r, rinv));
        } else if (g->mode == TGT_SPLIT) {
            // y_j = yh 2^30 + yl: y_j c = yh (2^30 c) + yl c (mod r), four terms between reductions.
            const double r = g->rf, rinv = g->rinv;
            double acc = 0.0;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                acc = acc + fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                acc = acc + fmulmod(yh[j], g->qhat30_f[j], g->qhat30_q[j], r);
                if (j & 1) acc = freduce(acc, r, rinv);
            }
            if (P.exact) acc = acc - fmulmod((double)e, g->qmod_f, g->qmod_q, r);
            acc = acc - g->hmod_f;
            o = d2u(fcanon(acc, r, rinv));
        } else if (g->mode == TGT_FP) {
            const double r = g->rf, rinv = g->rinv;
            double acc = 0.0;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                acc = acc + fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                if ((j & 3) == 3) acc = freduce(acc, r, rinv);
            }
            if (P.exact) acc = acc - fmulmod((double)e, g->qmod_f, g->qmod_q, r);
            acc = acc - g->hmod_f;
            o = d2u(fcanon(acc, r, rinv));
        } else {
            const PrimeConst C = pc_int(pcs + g->prime);
            if constexpr (MAXS <= 7) {
                // sum_j y_j c_j from 32-bit halves (y_j, c_j < 2^60): y c = y0 c0 + (y0 c1 +
                // y1 c0) 2^32 + y1 c1 2^64 with y1, c1 < 2^28, so the middle sum stays below
                // 7 * 2^61 < 2^64 and the high one below 2^59; one 32 x 32 -> 64 multiply per
                // partial product instead of full 64 x 64 products. The same exact sum.
                u64 l = 0, m = 0, h = 0;
                u32 carry = 0;
#pragma unroll
                for (u32 j = 0; j < MAXS; ++j) {
                    const u64 cj = g->qhat[j];
                    const u32 y0 = (u32)y[j], y1 = (u32)(y[j] >> 32), c0 = (u32)cj, c1 = (u32)(cj >> 32);
                    const u64 p = (u64)y0 * c0;
                    l += p;
                    carry += l < p ? 1u : 0u;
                    m += (u64)y0 * c1 + (u64)y1 * c0;
                    h += (u64)y1 * c1;
                }
                const u64 lo = l + (m << 32);
                const u64 hi = h + carry + (m >> 32) + (lo < l ? 1 : 0);
                o = reduce128(hi, lo, C);
            } else {
                Dot60 acc;
#pragma unroll
                for (u32 j = 0; j < MAXS; ++j) acc.add(y[j], g->qhat[j]);
                u64 hi, lo;
                acc.get(hi, lo);
                o = reduce128(hi, lo, C);
            }
            if (P.exact) o = sub_mod(o, shoup(e, g->qmod, g->qmod_s, C.q), C.q);
            o = sub_mod(o, g->hmod, C.q);
        }
        outp[(u64)g->limb * n + c] = o;
    }
}

}  // namespace ckh
