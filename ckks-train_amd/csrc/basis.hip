// Centered exact fast base conversion and the division finish, for gfx950: the vector kernel
// (its per-coefficient arithmetic in conv_arith.h, shared with the CPU replay) and the
// matrix-core kernel. FMA contraction is disabled for this file (the f64 overflow estimate must
// round like the CPU path).
#include "basis.h"

#pragma clang fp contract(off)

namespace ckh {

// One block: 256 coefficients of one job, targets [t0, t0 + tg) of its plan. YF: inputs in y form
// (a separate instantiation: the per-source product it skips would otherwise hold registers).
template <u32 MAXS, bool PIECES, bool YF>
__global__ __launch_bounds__(256) void k_conv(ConvTables T, const ConvJob* __restrict__ jobs, u32 tg) {
    const u32 n = 1u << T.log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const u32 t0 = blockIdx.z * tg;
    if (t0 >= P.ntgt || c >= n) return;
    conv_coeff<MAXS, PIECES>(T.src, T.tgt, T.pc, P, job.in, job.out, n, c, t0, min(t0 + tg, P.ntgt), YF);
}

// ---------------------------------------------------------------- matrix-core conversion

// The wave operations of the matrix-core conversion on gfx950.
struct MfWave {
    typedef mf_v4i v4i;
    typedef int v16i __attribute__((ext_vector_type(16)));
    static CKH_DEV v16i mfma(v4i a, v4i b, v16i d) { return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, d, 0, 0, 0); }
    static CKH_DEV double shfl_xor32(double v) { return __shfl_xor(v, 32); }
};

// One block: 256 coefficients of one job, all targets of its plan (conv_mf_block).
template <u32 KB, bool YF>
__global__ __launch_bounds__(256) void k_conv_mf(ConvTables T, const ConvJob* __restrict__ jobs) {
    __shared__ mf_v4i a_lds[MF_MAX_GROUPS * 2 * 64];
    __shared__ MfTgt t_lds[MF_MAX_GROUPS * 4];
    const ConvJob job = jobs[blockIdx.y];
    const ConvPlan P = T.plans[job.plan];
    const MfPlan F = T.mf.plans[job.plan];
    conv_mf_block<KB, MfWave>(T.src, P, F, T.mf.a, T.mf.tgt, job.in, job.out, 1u << T.log_n, blockIdx.x, threadIdx.x,
                              a_lds, t_lds, YF);
}

template <u32 MAXS, bool YF>
static void conv_launch_s(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 variant, u32 tg,
                          hipStream_t stream) {
    const u32 n = 1u << t.log_n;
    const u32 zt = (max_tgt + tg - 1) / tg;
    if (variant == 1)
        hipLaunchKernelGGL((k_conv<MAXS, true, YF>), dim3((n + 255) / 256, njobs, zt), dim3(256), 0, stream, t, jobs, tg);
    else
        hipLaunchKernelGGL((k_conv<MAXS, false, YF>), dim3((n + 255) / 256, njobs, zt), dim3(256), 0, stream, t, jobs, tg);
}

template <bool YF>
static void conv_launch_t(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                          u32 tg, hipStream_t stream) {
    const u32 n = 1u << t.log_n;
    if (variant == 2 && t.mf_ok && max_src <= 8 && n >= 256) {
        if (max_src <= 4)
            hipLaunchKernelGGL((k_conv_mf<1, YF>), dim3(n / 256, njobs), dim3(256), 0, stream, t, jobs);
        else
            hipLaunchKernelGGL((k_conv_mf<2, YF>), dim3(n / 256, njobs), dim3(256), 0, stream, t, jobs);
        return;
    }
    switch (max_src) {
        case 0:
        case 1: conv_launch_s<1, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 2: conv_launch_s<2, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 3: conv_launch_s<3, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 4: conv_launch_s<4, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 5: conv_launch_s<5, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 6: conv_launch_s<6, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 7: conv_launch_s<7, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        case 8: conv_launch_s<8, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
        default: conv_launch_s<16, YF>(t, jobs, njobs, max_tgt, variant, tg, stream); break;
    }
}

void conv_launch(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                 u32 tg, bool yform, hipStream_t stream) {
    if (njobs == 0 || max_tgt == 0) return;
    if (tg == 0) tg = 8;
    if (yform)
        conv_launch_t<true>(t, jobs, njobs, max_tgt, max_src, variant, tg, stream);
    else
        conv_launch_t<false>(t, jobs, njobs, max_tgt, max_src, variant, tg, stream);
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
