// NTT launchers (see ntt.hip for the schedule).
#pragma once
#include "common.h"

namespace ckh {

// Device tables shared by every transform of a context: per-prime constants and, per prime,
// N (w, w') pairs for the forward (psi^i at bitrev(i)) and inverse (psi^-i) directions.
struct NttTables {
    u32 log_n;
    const PrimeConst* pc;
    const Tw* fwd;     // [prime][N]
    const Tw* inv;     // [prime][N]
    const FTw* fwd_f;  // [prime][N], f64 path (primes below 2^50)
    const FTw* inv_f;
    // Row factors of the register-radix row pass (N = 2^(A+M), M = log N / 2 row stages):
    // rf[prime][row][s] = fwd_f[prime][2^(A+s) + row 2^s] for s < M (8 slots per row); ri the
    // same over inv_f. Every row-pass twiddle is a row factor times one of the first 2^M entries.
    const FTw* rf;
    const FTw* ri;
};

constexpr u32 ROW_FACTOR_SLOTS = 8;

// Register-radix passes (ntt_fast.hip) for 2^12 <= N <= 2^16.
bool ntt_fast_supported(u32 log_n);
void ntt_fast_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
void ntt_fast_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
// Forward NTT whose row pass ends with fin[r]: out = (a - NTT(src)) * inv (+ add) instead of
// storing the transform (fin is a FinishJob array, one per ref).
struct FinishJob;
void ntt_fast_forward_finish(const NttTables& t, const LimbRef* refs, const FinishJob* fin, u32 nrefs,
                             hipStream_t stream);

// ref.dst <- NTT(ref.src) for each of the nrefs descriptors (device array). src may equal
// dst (in place). Inputs reduced, outputs reduced.
void ntt_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
void ntt_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);

}  // namespace ckh
