// Centered exact RNS base conversion (ModUp, ModDown, rescale) and the division finish.
#pragma once
#include "common.h"
#include "conv_arith.h"

namespace ckh {

typedef int mf_v4i __attribute__((ext_vector_type(4)));
struct MfTables {
    const MfPlan* plans;  // indexed like ConvTables::plans
    const mf_v4i* a;      // A tiles
    const MfTgt* tgt;
};

// Conversion tables (device pointers).
struct ConvTables {
    const ConvSrc* src;
    const YForm* yf;  // the YForm of every source (same index as src)
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
// targets per workgroup of the vector kernels (the sources are read once per group of tg);
// yform: every job's input limbs are in y form (YForm), for kernels compiled for that input.
void conv_launch(const ConvTables& t, const ConvJob* jobs, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                 u32 tg, bool yform, hipStream_t stream);

void finish_launch(const FinishJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n,
                   hipStream_t stream);

}  // namespace ckh
