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
    u32 fp;                   // q_j < 2^50
    u32 prime;                // index into the context's primes
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
    // The correction terms as addends: nq = -Q mod r (e nq = -e Q mod r), nh = -(Q - 1)/2 mod r;
    // nq_c = nq centered in (-r/2, r/2]; nq_half / nh_half their 25-bit halves (r < 2^50).
    u64 nq, nh;
    double nq_c, nq_half[2], nh_half[2];
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
// yform: the input limbs already hold y_j = [x_j + h_j]_{q_j} (Q/q_j)^-1 mod q_j (YForm: an
// inverse NTT computed it in its last stage).
template <u32 MAXS, bool PIECES>
CKH_DEV void conv_coeff(const ConvSrc* srcs, const ConvTgt* tgts, const PrimeConst* pcs, const ConvPlan& P,
                        const u64* in_p, u64* out_p, u32 n, u32 c, u32 t0, u32 t1, bool yform) {
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
                if (yform) {
                    yd[j] = u2d(x);
                    y[j] = x;
                } else {
                    double sv = u2d(x) + (double)s->h;
                    sv = sv >= s->qf ? sv - s->qf : sv;
                    double v = fmulmod(sv, s->qhinv_f, s->qhinv_q, s->qf);
                    v = v < 0.0 ? v + s->qf : v;
                    yd[j] = v;
                    y[j] = d2u(v);
                }
            } else {
                y[j] = yform ? x : shoup(add_mod(x, s->h, s->q), s->qhinv, s->qhinv_s, s->q);
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
                // f64(y_j): yd already is (y_j < 2^50 converts exactly) for an f64 source
                const double prod = (s->fp ? yd[j] : u64_to_f64_rn(y[j])) * s->inv;
                f = f + prod;
            }
        }
    }
    const u64 e = P.exact ? (u64)f : 0;  // f >= 0: truncation is floor
    // With at most 8 sources, e < 8: -e (Q mod r) and -(Q - 1)/2 enter the sums as addends (e nq_c
    // and the half products are exact) instead of a modular product and two subtractions; every
    // f64 sum stays an integer below 2^53, which fcanon reduces exactly.
    const double ed = (double)e;
    const auto outp = glb(out_p);
    for (u32 t = t0; t < t1; ++t) {
        const auto* g = cst(tgts + P.tgt_first + t);
        u64 o;
        if (PIECES && g->mode == TGT_SPLIT && srcs_int) {
            // y_j = y0 + y1 2^20 + y2 2^40 (20-bit pieces), c_j = c0 + c1 2^25 (25-bit halves):
            // S_ab = sum_j y_ja c_jb is exact (every product < 2^45, at most 7 terms), and
            // sum_j y_j c_j = S00 + (32 S01 + S10) 2^20 + (32 S11 + S20) 2^40 + S21 2^65, the two
            // bracketed sums staying below 2^53.
            // -(Q - 1)/2 starts the two low sums and e (-Q mod r) is one more term of them (e < 7:
            // both stay below 7 2^45 + 2^28, the bracketed sums below 2^53)
            const double r = g->rf, rinv = g->rinv;
            double s00 = g->nh_half[0], s01 = g->nh_half[1], s10 = 0.0, s11 = 0.0, s20 = 0.0, s21 = 0.0;
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
            s00 = __builtin_fma(ed, g->nq_half[0], s00);
            s01 = __builtin_fma(ed, g->nq_half[1], s01);
            const double t1 = __builtin_fma(s01, 32.0, s10), t2 = __builtin_fma(s11, 32.0, s20);
            double acc = s00 + fmulmod(t1, g->p20.w, g->p20.wq, r);
            acc = acc + fmulmod(t2, g->p40.w, g->p40.wq, r);
            acc = acc + fmulmod(s21, g->p65.w, g->p65.wq, r);
            o = d2u(fcanon(acc, r, rinv));
        } else if (g->mode == TGT_SPLIT) {
            // y_j = yh 2^30 + yl: y_j c = yh (2^30 c) + yl c (mod r), four terms between reductions.
            const double r = g->rf, rinv = g->rinv;
            double acc = -g->hmod_f;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                acc = acc + fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                acc = acc + fmulmod(yh[j], g->qhat30_f[j], g->qhat30_q[j], r);
                if (j & 1) acc = freduce(acc, r, rinv);
            }
            if constexpr (MAXS <= 8)
                acc = __builtin_fma(ed, g->nq_c, acc);
            else if (P.exact)
                acc = acc - fmulmod(ed, g->qmod_f, g->qmod_q, r);
            o = d2u(fcanon(acc, r, rinv));
        } else if (g->mode == TGT_FP) {
            const double r = g->rf, rinv = g->rinv;
            double acc = -g->hmod_f;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                acc = acc + fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                if ((j & 3) == 3) acc = freduce(acc, r, rinv);
            }
            if constexpr (MAXS <= 8)
                acc = __builtin_fma(ed, g->nq_c, acc);
            else if (P.exact)
                acc = acc - fmulmod(ed, g->qmod_f, g->qmod_q, r);
            o = d2u(fcanon(acc, r, rinv));
        } else {
            const PrimeConst C = pc_int(pcs + g->prime);
            if constexpr (MAXS <= 7) {
                // sum_j y_j c_j from 32-bit halves (y_j, c_j < 2^60): y c = y0 c0 + (y0 c1 +
                // y1 c0) 2^32 + y1 c1 2^64 with y1, c1 < 2^28, so the middle sum stays below
                // 7 * 2^61 < 2^64 and the high one below 2^59; one 32 x 32 -> 64 multiply per
                // partial product instead of full 64 x 64 products. The same exact sum, which
                // starts at -(Q - 1)/2 mod r and takes e (-Q mod r) as one more term (e < 7).
                u64 l = g->nh, m = 0, h = 0;
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
                    // one term at a time: otherwise the compiler forms every term's products
                    // first and holds them all (92 VGPRs instead of 66 with y-form inputs)
                    CKH_OPAQUE(l);
                    CKH_OPAQUE(m);
                    CKH_OPAQUE(h);
                }
                {
                    const u64 p = (u64)(u32)e * (u32)g->nq;
                    l += p;
                    carry += l < p ? 1u : 0u;
                    m += (u64)(u32)e * (u32)(g->nq >> 32);
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
                if (P.exact) o = sub_mod(o, shoup(e, g->qmod, g->qmod_s, C.q), C.q);
                o = sub_mod(o, g->hmod, C.q);
            }
        }
        st_nt(outp, (u64)g->limb * n + c, o);
    }
}

// ---------------------------------------------------------------- matrix-core form
//
// Sum_j y_j c_jt computed as sum_p 2^(8p) sum_(j,b) y_j[b] w_p(2^(8b) c_jt mod r_t): y_j[b] are
// the bytes of y_j (offset by -128 to signed), w_p the balanced signed base-256 digits of the
// constant; one 32 x 32 x 32 int8 matrix product multiplies 32 (target, digit) rows by 32
// coefficients over 4 sources. Targets are grouped four at a time, f64-path and integer-path
// targets in separate groups.
constexpr u32 MF_MAX_GROUPS = 12;
struct MfPlan {
    u32 ngroups;  // 0: no matrix form (the vector kernel runs)
    u32 kb;       // K blocks of 32 (sources / 4, rounded up)
    u32 a_off;    // first A tile (64 lanes x 16 bytes) of the plan
    u32 t_off;    // first MfTgt of the plan (4 per group)
};
struct MfTgt {
    double rf, rinv, c32_f, c32_q, bias_f, nq_c;  // f64 path (nq_c: -Q mod r, centered)
    u64 q, nq, corr;                              // integer path (nq: -Q mod r; corr = bias_f)
    double nq_f, corr_f;                          // integer path: nq and corr as f64 (rounded)
    u32 limb;  // output limb; ~0u = padding
    u32 fp;    // 1: f64 recombination (r < 2^50)
};

// One block of the matrix-core conversion: 256 coefficients (block bx) of one job (plan P with
// matrix form F), all targets of the plan; thread tid of 256 (four waves of 64). W provides the
// wave operations: the operand types v4i (16 bytes per lane) and v16i (16 int32 per lane),
// mfma(a, b, d) (the wave's 32 x 32 x 32 int8 product: lane l holds A row l % 32 and B column
// l % 32 at k = 16 (l / 32) + byte, and receives D column l % 32, rows 8 (r / 4) + 4 (l / 32) +
// r % 4 in its element r) and shfl_xor32(v) (lane l ^ 32's value). a_lds / t_lds are the
// block's shared copies of the plan's A tiles and target constants.
// Each wave takes 2 tiles of
// 32 coefficients; lane l works on coefficient l % 32 of a tile and on sources 2h, 2h + 1,
// 4 + 2h, 5 + 2h (h = l / 32), which is exactly its B operand (k = 16 h + i: source 2h + i / 8,
// byte i % 8; second K block: sources + 4). The overflow estimate gathers the partner lane's
// products (lane l ^ 32) and sums them in source order.
template <u32 KB, class W>
CKH_DEV void conv_mf_block(const ConvSrc* srcs, const ConvPlan& P, const MfPlan& F, const typename W::v4i* a_tiles,
                           const MfTgt* tgts, const u64* in_p, u64* out_p, u32 n, u32 bx, u32 tid,
                           typename W::v4i* a_lds, MfTgt* t_lds, bool yform) {
    typedef typename W::v4i v4i;
    typedef typename W::v16i v16i;
    {
        const u32 na = F.ngroups * F.kb * 64;  // the plan's own K blocks (F.kb <= KB)
        for (u32 i = tid; i < na; i += 256) a_lds[i] = a_tiles[(u64)F.a_off * 64 + i];
        const u32 words = F.ngroups * 4 * (u32)(sizeof(MfTgt) / 8);
        const u64* src = (const u64*)(tgts + F.t_off);
        u64* dst = (u64*)t_lds;
        for (u32 i = tid; i < words; i += 256) dst[i] = src[i];
    }
    __syncthreads();
    const u32 lane = tid & 63, w = tid >> 6, h = lane >> 5, nl = lane & 31;
    const auto inp = glb(in_p);
    const auto outp = glb(out_p);
    constexpr u64 FLIP = 0x8080808080808080ull;  // byte - 128: the signed view of each byte
    // The wave's two tiles of 32 coefficients run side by side (two independent products per
    // group, one read of each target's constants for both).
    u32 c[2];
    v4i b[2][KB];
    u64 e[2] = {0, 0};
#pragma unroll
    for (u32 nt = 0; nt < 2; ++nt) {
        c[nt] = bx * 256 + w * 64 + nt * 32 + nl;
        u64 y[2 * KB];
        double pr[2 * KB];
#pragma unroll
        for (u32 q = 0; q < 2 * KB; ++q) {
            const u32 j = 4 * (q / 2) + 2 * h + (q % 2);
            y[q] = 0;
            pr[q] = 0.0;
            if (j < P.nsrc) {
                const auto* sp = cst(srcs + P.src_first + j);
                const u64 x = inp[(u64)(P.src_limb0 + j) * n + c[nt]];
                if (yform) {
                    y[q] = x;
                } else if (sp->fp) {
                    double sv = u2d(x) + (double)sp->h;
                    sv = sv >= sp->qf ? sv - sp->qf : sv;
                    double v = fmulmod(sv, sp->qhinv_f, sp->qhinv_q, sp->qf);
                    v = v < 0.0 ? v + sp->qf : v;
                    y[q] = d2u(v);
                } else {
                    y[q] = shoup(add_mod(x, sp->h, sp->q), sp->qhinv, sp->qhinv_s, sp->q);
                }
                pr[q] = u64_to_f64_rn(y[q]) * sp->inv;
            }
        }
        // e = floor(sum_j f64(y_j) / q_j) in source order j = 0, 1, ... (as the vector kernel)
        if (P.exact) {
            double other[2 * KB];
#pragma unroll
            for (u32 q = 0; q < 2 * KB; ++q) other[q] = W::shfl_xor32(pr[q]);
            double f = 0.0;
#pragma unroll
            for (u32 j = 0; j < 4 * KB; ++j) {
                if (j < P.nsrc) {
                    const u32 hj = (j / 2) % 2, q = 2 * (j / 4) + (j % 2);
                    f = f + (hj == h ? pr[q] : other[q]);
                }
            }
            e[nt] = (u64)f;
        }
#pragma unroll
        for (u32 kb = 0; kb < KB; ++kb) {
            const u64 y0 = y[2 * kb] ^ FLIP, y1 = y[2 * kb + 1] ^ FLIP;
            b[nt][kb] = v4i{(int)(u32)y0, (int)(u32)(y0 >> 32), (int)(u32)y1, (int)(u32)(y1 >> 32)};
        }
    }
    auto product = [&](u32 g, u32 nt) {
        v16i d = {};
#pragma unroll
        for (u32 kb = 0; kb < KB; ++kb)
            if (kb < F.kb)
                d = W::mfma(a_lds[(g * F.kb + kb) * 64 + lane], b[nt][kb], d);
        return d;
    };
    v16i dn0 = product(0, 0), dn1 = product(0, 1);
#pragma unroll 1
    for (u32 g = 0; g < F.ngroups; ++g) {
        const v16i d0 = dn0, d1 = dn1;
        if (g + 1 < F.ngroups) {  // the next group's products overlap this group's recombination
            dn0 = product(g + 1, 0);
            dn1 = product(g + 1, 1);
        }
#pragma unroll
        for (u32 uu = 0; uu < 2; ++uu) {
            const MfTgt& tg = t_lds[4 * g + 2 * h + uu];
            const u32 limb = tg.limb;
            if (limb == ~0u) continue;
            const bool fp = tg.fp;
            const double r = tg.rf, rinv = tg.rinv, c32_f = tg.c32_f, c32_q = tg.c32_q, bias_f = tg.bias_f;
            const double nq_c = tg.nq_c;
#pragma unroll
            for (u32 nt = 0; nt < 2; ++nt) {
                const v16i& d = nt ? d1 : d0;
                // X = lo + hi 2^32, lo / hi = sum of four digit planes each (|.| < 2^45, exact);
                // planes p and p + 1 combine in int32 first (|d| < 2^20: |d_p + 256 d_(p+1)| < 2^29)
                const int l01 = d[8 * uu + 0] + d[8 * uu + 1] * 256, l23 = d[8 * uu + 2] + d[8 * uu + 3] * 256;
                const int h01 = d[8 * uu + 4] + d[8 * uu + 5] * 256, h23 = d[8 * uu + 6] + d[8 * uu + 7] * 256;
                const double lo = __builtin_fma((double)l23, 65536.0, (double)l01);
                const double hi = __builtin_fma((double)h23, 65536.0, (double)h01);
                u64 o;
                if (fp) {
                    // e < 8: e nq_c is exact and the sum stays an integer below 2^53
                    double acc = lo + fmulmod(hi, c32_f, c32_q, r) + bias_f;
                    acc = __builtin_fma((double)e[nt], nq_c, acc);
                    o = d2u(fcanon(acc, r, rinv));
                } else {
                    // V = X + A with X = lo + hi 2^32 (|lo|, |hi| < 2^45) and A = e nq + corr
                    // (below 2^64). k = rint(V / r) from f64 estimates: their error (below 2^27 on
                    // |V| < 2^79, then the reciprocal's) moves V / r by less than 2^-30, so
                    // |V - k r| < r; V - k r is then computed exactly in wrapping 64-bit arithmetic
                    // and one addition of r makes it non-negative.
                    const double T = __builtin_fma(hi, 4294967296.0, lo) +
                                     __builtin_fma((double)e[nt], tg.nq_f, tg.corr_f);
                    const int64_t k = (int64_t)__builtin_rint(T * rinv);
                    const u64 lo64 = (u64)((int64_t)l01 + (int64_t)l23 * 65536);
                    const u64 hi64 = (u64)((int64_t)h01 + (int64_t)h23 * 65536);
                    const u32 ee = (u32)e[nt];
                    const u64 A = (u64)ee * (u32)tg.nq + (((u64)ee * (u32)(tg.nq >> 32)) << 32) + tg.corr;
                    const u64 R = lo64 + (hi64 << 32) + A - (u64)k * tg.q;
                    o = (int64_t)R < 0 ? R + tg.q : R;
                }
                outp[(u64)limb * n + c[nt]] = o;
            }
        }
    }
}

}  // namespace ckh
