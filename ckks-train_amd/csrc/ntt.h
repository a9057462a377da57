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


// Register-radix passes (ntt_fast.hip) for 2^12 <= N <= 2^16.
bool ntt_fast_supported(u32 log_n);

// The one-launch forward transform (both passes in one persistent launch, the intermediate handed
// from the column pass to the row pass through the L2 of one XCD). Its control block: a ticket
// counter per XCD, CTL_STRIDE words apart, then one done-counter per limb; zero before each
// launch (fused_ctl_words(nrefs) u32 words).
constexpr u32 CTL_STRIDE = 64;
constexpr u32 FUSED_CTL_HEAD = 8 * CTL_STRIDE;
inline u64 fused_ctl_words(u32 nrefs) { return FUSED_CTL_HEAD + nrefs; }
struct FusedCfg {
    u32 nx = 0;      // XCDs (0: the one-launch transform is unavailable)
    u32 lag = 8;     // slots between a limb's first-pass tiles and its second-pass tiles
    u32 grid = 0;    // persistent workgroups
    u32* err = nullptr;  // device word set when a wait overran its bound
};
// XCD count of this device: the workgroups of a probe grid report XCD ids 0..count-1, each id on
// a fair share of them (else 0: the one-launch transform stays off).
u32 xcc_layout_check(hipStream_t stream);
// Workgroups of the one-launch transform that fit on one CU at once.
u32 fused_occupancy();
struct FinishJob;
// ntt_fast_forward (fin null) or ntt_fast_forward_finish in one launch; ctl zeroed, nrefs words
// of done-counters after its head.
void ntt_fast_forward_fused(const NttTables& t, const LimbRef* refs, u32 nrefs, const FinishJob* fin, u32* ctl,
                            const FusedCfg& f, hipStream_t stream);
// ntt_fast_inverse in one launch (row pass, then the column pass in place on dst).
void ntt_fast_inverse_fused(const NttTables& t, const LimbRef* refs, u32 nrefs, u32* ctl, const FusedCfg& f,
                            hipStream_t stream);
void ntt_fast_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
// yf non-null: every ref leaves in the y form of the conversion source yf[ref.aux] (YForm).
void ntt_fast_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream,
                      const YForm* yf = nullptr);
// Forward NTT whose row pass ends with fin[r]: out = (a - NTT(src)) * inv (+ add) instead of
// storing the transform (fin is a FinishJob array, one per ref).
void ntt_fast_forward_finish(const NttTables& t, const LimbRef* refs, const FinishJob* fin, u32 nrefs,
                             hipStream_t stream);

// The two passes of ntt_fast_forward separately (column pass: src -> dst; row pass: in place on
// dst; between them dst holds f64 intermediates, not residues).
void ntt_fast_forward_cols(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
void ntt_fast_forward_rows(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);

void rows_ip_launch(const NttTables& t, const RowIpJob* jobs, u32 njobs, hipStream_t stream);

void rows_ip_fin_launch(const NttTables& t, const RowIpFinJob* jobs, u32 njobs, hipStream_t stream);

// ref.dst <- NTT(ref.src) for each of the nrefs descriptors (device array). src may equal
// dst (in place). Inputs reduced, outputs reduced.
void ntt_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);
void ntt_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream);

}  // namespace ckh
