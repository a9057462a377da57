// Centered exact RNS base conversion (ModUp, ModDown, rescale) and the division finish.
#pragma once
#include "common.h"
#include "conv_arith.h"

namespace ckh {

// Matrix-core (int8 MFMA) form of a conversion plan. Sum_j y_j c_jt is computed as
// sum_p 2^(8p) sum_(j,b) y_j[b] w_p(2^(8b) c_jt mod r_t): y_j[b] are the bytes of y_j (offset by
// -128 to signed), w_p the balanced signed base-256 digits of the constant; one
// v_mfma_i32_32x32x32_i8 multiplies 32 (target, digit) rows by 32 coefficients over 4 sources.
// Targets are grouped four at a time, f64-path and integer-path targets in separate groups.
constexpr u32 MF_MAX_GROUPS = 12;
struct MfPlan {
    u32 ngroups;  // 0: no matrix form (the vector kernel runs)
    u32 kb;       // K blocks of 32 (sources / 4, rounded up)
    u32 a_off;    // first A tile (64 lanes x 16 bytes) of the plan
    u32 t_off;    // first MfTgt of the plan (4 per group)
};
struct MfTgt {
    double rf, rinv, c32_f, c32_q, bias_f, qmod_f, qmod_q;  // f64 path
    u64 q, one_s, two64, two64_s, qmod, qmod_s, corr;       // integer path
    u32 limb;  // output limb; ~0u = padding
    u32 fp;    // 1: f64 recombination (r < 2^50)
};
typedef int mf_v4i __attribute__((ext_vector_type(4)));
struct MfTables {
    const MfPlan* plans;  // indexed like ConvTables::plans
    const mf_v4i* a;      // A tiles
    const MfTgt* tgt;
};

// Conversion tables (device pointers).
struct ConvTables {
    const ConvSrc* src;
    const ConvTgt* tgt;
    const ConvPlan* plans;
    const PrimeConst* pc;
    u32 log_n;
    MfTables mf;
    u32 mf_ok;  // every plan has a matrix form (variant 2 may run)
};

// For each job: coefficient-form input limbs (reduced) -> converted output limbs (reduced,
// coefficient form): out_t = [x_c]_r with x_c the centered value of the source residues
// (BaseConverter::convert_exact of [x + h]_Q minus h mod r).
// Tuning choices: variant: 0 = vector kernel; 1 = vector kernel with 50-bit targets of 60-bit
// sources summed as exact f64 pieces (20-bit source pieces times 25-bit constant halves) instead
// of two modular products per source; 2 = the matrix-core kernel (when ConvTables::mf_ok; else
// variant 0); tg:
// targets per workgroup of the vector kernels (the sources are read once per group of tg).
void conv_launch(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                 u32 tg, hipStream_t stream);

void finish_launch(const FinishJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n,
                   hipStream_t stream);

}  // namespace ckh
