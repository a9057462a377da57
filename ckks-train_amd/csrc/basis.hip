// Centered exact fast base conversion and the division finish, for gfx950: the vector kernel
// (its per-coefficient arithmetic in conv_arith.h, shared with the CPU replay) and the
// matrix-core kernel. FMA contraction is disabled for this file (the f64 overflow estimate must
// round like the CPU path).
#include "basis.h"

#pragma clang fp contract(off)

namespace ckh {

// One block: 256 coefficients of one job, targets [t0, t0 + tg) of its plan.
template <u32 MAXS, bool PIECES>
__global__ __launch_bounds__(256) void k_conv(ConvTables T, const ConvJob* __restrict__ jobs, u32 tg) {
    const u32 n = 1u << T.log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const u32 t0 = blockIdx.z * tg;
    if (t0 >= P.ntgt || c >= n) return;
    conv_coeff<MAXS, PIECES>(T.src, T.tgt, T.pc, P, job.in, job.out, n, c, t0, min(t0 + tg, P.ntgt));
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
