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

// f64 helpers for primes below 2^50 (exactness argument in ntt_fast.hip): u2d/d2u convert
// exactly, fmulmod returns y w mod r as an integer in [-0.75 r, 0.75 r] for 0 <= y < 2^50.
CKH_DEV double cv_u2d(u64 v) { return __builtin_fma((double)(u32)(v >> 32), 4294967296.0, (double)(u32)v); }

CKH_DEV u64 cv_d2u(double d) {
    const u32 hi = (u32)(d * (1.0 / 4294967296.0));
    const double lo = __builtin_fma(-(double)hi, 4294967296.0, d);
    return ((u64)hi << 32) | (u32)lo;
}

CKH_DEV double cv_fmulmod(double y, double w, double wq, double r) {
    const double h = y * w;
    const double l = __builtin_fma(y, w, -h);
    const double k = __builtin_rint(y * wq);
    return __builtin_fma(-k, r, h) + l;
}

template <u32 MAXS>
__global__ __launch_bounds__(256) void k_conv(ConvTables T, const ConvJob* __restrict__ jobs, u32 tg) {
    const u32 n = 1u << T.log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const u32 t0 = blockIdx.z * tg;
    if (t0 >= P.ntgt || c >= n) return;
    const u32 t1 = min(t0 + tg, P.ntgt);
    const u64* inp = job.in;
    u64 y[MAXS];
    double yd[MAXS], yh[MAXS];
    double f = 0.0;
#pragma unroll
    for (u32 j = 0; j < MAXS; ++j) {
        y[j] = 0;
        yd[j] = 0.0;
        yh[j] = 0.0;
        if (j < P.nsrc) {
            const ConvSrc* s = T.src + P.src_first + j;
            const u64 x = inp[(u64)(P.src_limb0 + j) * n + c];
            if (s->fp) {
                double sv = cv_u2d(x) + (double)s->h;
                sv = sv >= s->qf ? sv - s->qf : sv;
                double v = cv_fmulmod(sv, s->qhinv_f, s->qhinv_q, s->qf);
                v = v < 0.0 ? v + s->qf : v;
                yd[j] = v;
                y[j] = cv_d2u(v);
            } else {
                y[j] = shoup(add_mod(x, s->h, s->q), s->qhinv, s->qhinv_s, s->q);
                yd[j] = (double)(u32)(y[j] & ((1u << 30) - 1));  // low 30 bits, exact
                yh[j] = (double)(u32)(y[j] >> 30);                // high 30 bits, exact
            }
            if (P.exact) {
                const double prod = u64_to_f64_rn(y[j]) * s->inv;
                f = f + prod;
            }
        }
    }
    const u64 e = P.exact ? (u64)f : 0;  // f >= 0: truncation is floor
    u64* outp = job.out;
    for (u32 t = t0; t < t1; ++t) {
        const ConvTgt* g = T.tgt + P.tgt_first + t;
        u64 o;
        if (g->mode == TGT_SPLIT) {
            // y_j = yh 2^30 + yl: y_j c = yh (2^30 c) + yl c (mod r), four terms between reductions.
            const double r = g->rf, rinv = g->rinv;
            double acc = 0.0;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                if (j < P.nsrc) {
                    acc = acc + cv_fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                    acc = acc + cv_fmulmod(yh[j], g->qhat30_f[j], g->qhat30_q[j], r);
                    if (j & 1) acc = __builtin_fma(-__builtin_rint(acc * rinv), r, acc);
                }
            }
            if (P.exact) acc = acc - cv_fmulmod((double)e, g->qmod_f, g->qmod_q, r);
            acc = acc - g->hmod_f;
            acc = __builtin_fma(-__builtin_rint(acc * rinv), r, acc);
            o = cv_d2u(acc < 0.0 ? acc + r : acc);
        } else if (g->mode == TGT_FP) {
            const double r = g->rf, rinv = g->rinv;
            double acc = 0.0;
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j) {
                if (j < P.nsrc) {
                    acc = acc + cv_fmulmod(yd[j], g->qhat_f[j], g->qhat_q[j], r);
                    if ((j & 3) == 3) acc = __builtin_fma(-__builtin_rint(acc * rinv), r, acc);
                }
            }
            if (P.exact) acc = acc - cv_fmulmod((double)e, g->qmod_f, g->qmod_q, r);
            acc = acc - g->hmod_f;
            acc = __builtin_fma(-__builtin_rint(acc * rinv), r, acc);
            o = cv_d2u(acc < 0.0 ? acc + r : acc);
        } else {
            const PrimeConst C = T.pc[g->prime];
            Acc128 acc = {0, 0};
#pragma unroll
            for (u32 j = 0; j < MAXS; ++j)
                if (j < P.nsrc) mac(acc, y[j], g->qhat[j]);
            o = reduce128(acc.hi, acc.lo, C);
            if (P.exact) o = sub_mod(o, shoup(e, g->qmod, g->qmod_s, C.q), C.q);
            o = sub_mod(o, g->hmod, C.q);
        }
        outp[(u64)g->limb * n + c] = o;
    }
}

void conv_launch(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src,
                 hipStream_t stream) {
    if (njobs == 0 || max_tgt == 0) return;
    const u32 n = 1u << t.log_n;
    const u32 tg = 8;
    dim3 grid((n + 255) / 256, njobs, (max_tgt + tg - 1) / tg);
    if (max_src <= 1)
        hipLaunchKernelGGL(k_conv<1>, grid, dim3(256), 0, stream, t, jobs, tg);
    else if (max_src <= 4)
        hipLaunchKernelGGL(k_conv<4>, grid, dim3(256), 0, stream, t, jobs, tg);
    else if (max_src <= 8)
        hipLaunchKernelGGL(k_conv<8>, grid, dim3(256), 0, stream, t, jobs, tg);
    else
        hipLaunchKernelGGL(k_conv<16>, grid, dim3(256), 0, stream, t, jobs, tg);
}

__global__ __launch_bounds__(256) void k_finish(const FinishJob* __restrict__ jobs,
                                                const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const FinishJob job = jobs[blockIdx.y];
    const u64 q = pc[job.prime].q;
    u64 v = shoup(sub_mod(job.a[c], job.t[c], q), job.inv, job.inv_s, q);
    if (job.add) v = add_mod(v, job.add[c], q);
    job.out[c] = v;
}

void finish_launch(const FinishJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n, hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_finish, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, pc, log_n);
}

}  // namespace ckh
