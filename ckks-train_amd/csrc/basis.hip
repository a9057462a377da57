// Centered exact fast base conversion and the division finish, for gfx950.
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
// order from +0.0, then floor); FMA contraction is disabled for this file.
#include "basis.h"

#pragma clang fp contract(off)

namespace ckh {

// Round-to-nearest u64 -> f64: both halves convert exactly and the sum rounds once.
CKH_DEV double u64_to_f64_rn(u64 v) {
    const double hi = (double)(u32)(v >> 32) * 4294967296.0;
    const double lo = (double)(u32)v;
    return hi + lo;
}

// One block: 256 coefficients of one job, targets [t0, t0 + tg) of its plan; per-target
// constants are read through the scalar cache. 60-bit targets sum exact 128-bit integer
// products and reduce once.
template <u32 MAXS, bool PIECES>
__global__ __launch_bounds__(256) void k_conv(ConvTables T, const ConvJob* __restrict__ jobs, u32 tg) {
    const u32 n = 1u << T.log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const u32 t0 = blockIdx.z * tg;
    if (t0 >= P.ntgt || c >= n) return;
    const u32 t1 = min(t0 + tg, P.ntgt);
    const auto inp = glb(job.in);
    u64 y[MAXS];
    // f64 source: yd = y, yh = 0; integer source: yd / yh = its low / high 30 bits, or (PIECES)
    // yd / yh / yx = its 20-bit pieces from the low end
    double yd[MAXS], yh[MAXS], yx[MAXS];
    double f = 0.0;
    // the piece path needs every source integer (at most 7: the bracketed sums below 2^53)
    bool srcs_int = PIECES && MAXS <= 7;
    for (u32 j = 0; j < P.nsrc && srcs_int; ++j) srcs_int = !cst(T.src + P.src_first + j)->fp;
#pragma unroll
    for (u32 j = 0; j < MAXS; ++j) {
        y[j] = 0;
        yd[j] = 0.0;
        yh[j] = 0.0;
        yx[j] = 0.0;
        if (j < P.nsrc) {
            const ConvSrc* s = T.src + P.src_first + j;
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
    const auto outp = glb(job.out);
    for (u32 t = t0; t < t1; ++t) {
        const auto* g = cst(T.tgt + P.tgt_first + t);
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
            const PrimeConst C = pc_int(T.pc + g->prime);
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

// ---------------------------------------------------------------- matrix-core conversion
//
// One block: 256 coefficients of one job, all targets of its plan. Each wave takes 2 tiles of
// 32 coefficients; lane l works on coefficient l % 32 of a tile and on sources 2h, 2h + 1,
// 4 + 2h, 5 + 2h (h = l / 32), which is exactly its B operand (k = 16 h + i: source 2h + i / 8,
// byte i % 8; second K block: sources + 4). The overflow estimate gathers the partner lane's
// products (lane l ^ 32) and sums them in source order.
template <u32 KB>
__global__ __launch_bounds__(256) void k_conv_mf(ConvTables T, const ConvJob* __restrict__ jobs) {
    __shared__ mf_v4i a_lds[MF_MAX_GROUPS * 2 * 64];
    __shared__ MfTgt t_lds[MF_MAX_GROUPS * 4];
    const u32 n = 1u << T.log_n;
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const MfPlan F = T.mf.plans[job.plan];
    const u32 tid = threadIdx.x;
    {
        const u32 na = F.ngroups * F.kb * 64;  // the plan's own K blocks (F.kb <= KB)
        for (u32 i = tid; i < na; i += 256) a_lds[i] = T.mf.a[(u64)F.a_off * 64 + i];
        const u32 words = F.ngroups * 4 * (u32)(sizeof(MfTgt) / 8);
        const u64* src = (const u64*)(T.mf.tgt + F.t_off);
        u64* dst = (u64*)t_lds;
        for (u32 i = tid; i < words; i += 256) dst[i] = src[i];
    }
    __syncthreads();
    const u32 lane = tid & 63, w = tid >> 6, h = lane >> 5, nl = lane & 31;
    const auto inp = glb(job.in);
    const auto outp = glb(job.out);
    constexpr u64 FLIP = 0x8080808080808080ull;  // byte - 128: the signed view of each byte
    typedef int v16i __attribute__((ext_vector_type(16)));
    // The wave's two tiles of 32 coefficients run side by side (two independent products per
    // group, one read of each target's constants for both).
    u32 c[2];
    mf_v4i b[2][KB];
    u64 e[2] = {0, 0};
#pragma unroll
    for (u32 nt = 0; nt < 2; ++nt) {
        c[nt] = blockIdx.x * 256 + w * 64 + nt * 32 + nl;
        u64 y[2 * KB];
        double pr[2 * KB];
#pragma unroll
        for (u32 q = 0; q < 2 * KB; ++q) {
            const u32 j = 4 * (q / 2) + 2 * h + (q % 2);
            y[q] = 0;
            pr[q] = 0.0;
            if (j < P.nsrc) {
                const auto* sp = cst(T.src + P.src_first + j);
                const u64 x = inp[(u64)(P.src_limb0 + j) * n + c[nt]];
                if (sp->fp) {
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
            for (u32 q = 0; q < 2 * KB; ++q) other[q] = __shfl_xor(pr[q], 32);
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
            b[nt][kb] = mf_v4i{(int)(u32)y0, (int)(u32)(y0 >> 32), (int)(u32)y1, (int)(u32)(y1 >> 32)};
        }
    }
    auto product = [&](u32 g, u32 nt) {
        v16i d = {};
#pragma unroll
        for (u32 kb = 0; kb < KB; ++kb)
            if (kb < F.kb)
                d = __builtin_amdgcn_mfma_i32_32x32x32_i8(a_lds[(g * F.kb + kb) * 64 + lane], b[nt][kb], d, 0, 0, 0);
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
            const double qmod_f = tg.qmod_f, qmod_q = tg.qmod_q;
#pragma unroll
            for (u32 nt = 0; nt < 2; ++nt) {
                const v16i& d = nt ? d1 : d0;
                // X = lo + hi 2^32, lo / hi = sum of four digit planes each (|.| < 2^45, exact)
                double lo = (double)d[8 * uu + 3];
                lo = __builtin_fma(lo, 256.0, (double)d[8 * uu + 2]);
                lo = __builtin_fma(lo, 256.0, (double)d[8 * uu + 1]);
                lo = __builtin_fma(lo, 256.0, (double)d[8 * uu + 0]);
                double hi = (double)d[8 * uu + 7];
                hi = __builtin_fma(hi, 256.0, (double)d[8 * uu + 6]);
                hi = __builtin_fma(hi, 256.0, (double)d[8 * uu + 5]);
                hi = __builtin_fma(hi, 256.0, (double)d[8 * uu + 4]);
                u64 o;
                if (fp) {
                    double acc = lo + fmulmod(hi, c32_f, c32_q, r) + bias_f;
                    if (P.exact) acc = acc - fmulmod((double)e[nt], qmod_f, qmod_q, r);
                    o = d2u(fcanon(acc, r, rinv));
                } else {
                    // X + 2^45 + 2^77 = (lo + 2^45) + (hi + 2^45) 2^32 >= 0, below 2^79
                    const u64 lu = d2u(lo + 0x1p45), hu = d2u(hi + 0x1p45);
                    const u64 xl = lu + (hu << 32);
                    const u64 xh = (hu >> 32) + (xl < lu ? 1 : 0);
                    PrimeConst C{};
                    C.q = tg.q;
                    C.one_s = tg.one_s;
                    C.two64 = tg.two64;
                    C.two64_s = tg.two64_s;
                    o = add_mod(reduce128(xh, xl, C), tg.corr, C.q);
                    if (P.exact) o = sub_mod(o, shoup(e[nt], tg.qmod, tg.qmod_s, C.q), C.q);
                }
                outp[(u64)limb * n + c[nt]] = o;
            }
        }
    }
}

template <u32 MAXS>
static void conv_launch_s(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 variant, u32 tg,
                          hipStream_t stream) {
    const u32 n = 1u << t.log_n;
    const u32 zt = (max_tgt + tg - 1) / tg;
    if (variant == 1)
        hipLaunchKernelGGL((k_conv<MAXS, true>), dim3((n + 255) / 256, njobs, zt), dim3(256), 0, stream, t, jobs, tg);
    else
        hipLaunchKernelGGL((k_conv<MAXS, false>), dim3((n + 255) / 256, njobs, zt), dim3(256), 0, stream, t, jobs, tg);
}

void conv_launch(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                 u32 tg, hipStream_t stream) {
    if (njobs == 0 || max_tgt == 0) return;
    if (tg == 0) tg = 8;
    const u32 n = 1u << t.log_n;
    if (variant == 2 && t.mf_ok && max_src <= 8 && n >= 256) {
        if (max_src <= 4)
            hipLaunchKernelGGL(k_conv_mf<1>, dim3(n / 256, njobs), dim3(256), 0, stream, t, jobs);
        else
            hipLaunchKernelGGL(k_conv_mf<2>, dim3(n / 256, njobs), dim3(256), 0, stream, t, jobs);
        return;
    }
    switch (max_src) {
        case 0:
        case 1: conv_launch_s<1>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 2: conv_launch_s<2>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 3: conv_launch_s<3>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 4: conv_launch_s<4>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 5: conv_launch_s<5>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 6: conv_launch_s<6>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 7: conv_launch_s<7>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 8: conv_launch_s<8>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        default: conv_launch_s<16>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
    }
}

__global__ __launch_bounds__(256) void k_finish(const FinishJob* __restrict__ jobs,
                                                const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const FinishJob job = jobs[blockIdx.y];
    const u64 q = pc[job.prime].q;
    u64 v = shoup(sub_mod(glb(job.a)[c], glb(job.t)[c], q), job.inv, job.inv_s, q);
    if (job.add) v = add_mod(v, glb(job.add)[c], q);
    glb(job.out)[c] = v;
}

void finish_launch(const FinishJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n, hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_finish, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, pc, log_n);
}

}  // namespace ckh
