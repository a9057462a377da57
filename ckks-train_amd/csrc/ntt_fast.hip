// Register-radix NTT kernels for 2^12 <= N <= 2^16 (gfx950): the passes of ntt_passes.h as
// kernels over batches of limbs, the row pass fused with the key inner product, and the
// one-launch forms that hand a limb from its first pass to its second inside one launch.
#include <algorithm>
#include <vector>

#include "basis.h"
#include "common.h"
#include "ntt.h"

namespace ckh {
namespace fast {

// A word of a column-pass output written by another workgroup of the same launch: loaded at agent
// scope (the sc1 bit), so the load misses this CU's vector L1 (which may hold an older copy of the
// line) and reads the XCD's L2, where the writer's completed stores are.
CKH_DEV u64 load_coherent(const u64* p) {
    return __hip_atomic_load(glb(const_cast<u64*>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace fast
}  // namespace ckh

#include "ntt_passes.h"

namespace ckh {
namespace fast {

// s_getreg operand for HW_REG_XCC_ID (register 20), bits 0..3: the XCD the wave runs on.
constexpr int XCC_ID_HWREG = 20 | (0 << 6) | ((4 - 1) << 11);
// s_waitcnt operand: vmcnt(0), expcnt and lgkmcnt not waited for.
constexpr int WAITCNT_VM0 = 0x0F70;
constexpr u32 FUSED_SPIN_MAX = 1u << 22;

// ---------------------------------------------------------------- kernels

// Block -> (limb, tile): workgroups are dealt round-robin to the 8 XCDs by linear id, so the
// tiles of one limb are given ids that land on one XCD (limbs in groups of 8); the column and
// row passes of a limb then run on the same XCD and the intermediate can stay in its L2.
CKH_DEV void xcd_map(u32 b, u32 tiles, u32 nrefs, u32& ref, u32& tile) {
    const u32 full = (nrefs / 8) * 8 * tiles;
    if (b < full) {
        const u32 xcd = b & 7, slot = b >> 3;
        ref = (slot / tiles) * 8 + xcd;
        tile = slot % tiles;
    } else {
        ref = (nrefs / 8) * 8 + (b - full) / tiles;
        tile = (b - full) % tiles;
    }
}

#define CKH_AR_DISPATCH(CALL)                                                                   \
    u32 ref_i, tile;                                                                            \
    xcd_map(blockIdx.x, 1u << (log_n - TL), nrefs, ref_i, tile);                                \
    const LimbRef ref = refs[ref_i];                                                            \
    const PrimeConst* pc = pcs + ref.prime;                                                     \
    const u64 toff = (u64)ref.prime << log_n;                                                   \
    if (pc->q < FP_PRIME_BOUND) {                                                               \
        const FpAr ar{twf + toff, pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f, rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};                \
        CALL;                                                                                   \
    } else {                                                                                    \
        const IntAr ar{tw + toff, pc->q, pc->n_inv, pc->n_inv_s, pc->w_n_inv, pc->w_n_inv_s};   \
        CALL;                                                                                   \
    }

template <u32 M>
__global__ __launch_bounds__(NT) void k_fwd_cols(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                 const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                 const FTw* __restrict__ rt, u32 log_n,
                                                 u32 nrefs) {
    __shared__ u64 lds[LDS_WORDS];
    CKH_AR_DISPATCH((fwd_cols<M>(ar, ref.src, ref.dst, lds, log_n, tile)))
}

// FIN: the finish variant (a separate instantiation: its prefetched operands cost registers).
template <u32 M, bool FIN>
__global__ __launch_bounds__(NT) void k_fwd_rows(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                 const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                 const FTw* __restrict__ rt, u32 log_n,
                                                 u32 nrefs, const FinishJob* __restrict__ fins) {
    __shared__ u64 lds[LDS_WORDS];
    CKH_AR_DISPATCH((fwd_rows<M>(ar, ref.dst, ref.dst, lds, log_n, FIN ? fins + ref_i : nullptr, pc->q, tile)))
}

template <u32 M>
__global__ __launch_bounds__(NT) void k_inv_rows(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                 const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                 const FTw* __restrict__ rt, u32 log_n,
                                                 u32 nrefs) {
    __shared__ u64 lds[LDS_WORDS];
    CKH_AR_DISPATCH((inv_rows<M>(ar, ref.src, ref.dst, lds, log_n, tile)))
}

// YF: every limb leaves in the y form of the conversion source yf[ref.aux] (YForm).
template <u32 M, bool YF>
__global__ __launch_bounds__(NT) void k_inv_cols(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                 const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                 const FTw* __restrict__ rt, u32 log_n,
                                                 u32 nrefs, const YForm* __restrict__ yf) {
    __shared__ u64 lds[LDS_WORDS];
    if constexpr (!YF) {
        CKH_AR_DISPATCH((inv_cols<M>(ar, ref.dst, lds, log_n, tile)))
    } else {
        u32 ref_i, tile;
        xcd_map(blockIdx.x, 1u << (log_n - TL), nrefs, ref_i, tile);
        const LimbRef ref = refs[ref_i];
        const PrimeConst* pc = pcs + ref.prime;
        const auto* y = cst(yf + ref.aux);
        const u64 toff = (u64)ref.prime << log_n;
        if (pc->q < FP_PRIME_BOUND) {
            const FTw nf{y->n_f.w, y->n_f.wq}, wnf{y->wn_f.w, y->wn_f.wq};
            FpAr ar{twf + toff, pc->qf, pc->qinv, nf, wnf, rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
            ar.add = y->h_f;
            inv_cols<M, FpAr, false, true>(ar, ref.dst, lds, log_n, tile);
        } else {
            IntAr ar{tw + toff, pc->q, y->n, y->n_s, y->wn, y->wn_s};
            ar.add = y->h;
            inv_cols<M, IntAr, false, true>(ar, ref.dst, lds, log_n, tile);
        }
    }
}

// ---------------------------------------------------------------- one-launch forward NTT
//
// Both passes of the forward transform in one persistent launch, so that a limb's column-pass
// output is read back by its row pass from the L2 of the XCD that wrote it instead of HBM. The
// limbs are split into one contiguous chunk per XCD (nx of them, ids from the hardware register;
// checked at context creation); the workgroups running on an XCD take tickets from that XCD's
// counter in the order they start, and ticket t (slot s = t / tiles, tile w = t % tiles) runs
// column tile w of the chunk's limb s, then row tile w of its limb s - lag. A row tile waits
// (thread 0, bounded) until every column tile of its limb has been counted done: those tiles hold
// earlier tickets of the same counter, so their workgroups are running or finished and the wait
// ends. Column tiles count themselves done after their stores have completed (vmcnt 0) and row
// tiles read the intermediate with agent-scope loads (load_coherent); the limb's words pass
// through the L2 of that XCD only, which every one of its CUs reads and writes. An overlong wait
// sets the error word (the row tile then runs anyway; the host reports the error).
//
// Control block (u32 words, zero before the launch): ticket counter of XCD x at x * CTL_STRIDE,
// then one done-counter per limb from FUSED_CTL_HEAD. err: the context's device error word.
template <u32 MA, u32 MB, bool FIN>
__global__ __launch_bounds__(NT) void k_fwd_fused(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                  const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                  const FTw* __restrict__ rt, u32 log_n, u32 nrefs,
                                                  const FinishJob* __restrict__ fins, u32* ctl, u32 nx, u32 lag,
                                                  u32* err) {
    __shared__ u64 lds[LDS_WORDS];
    __shared__ u32 s_ticket;
    const u32 xcc = __builtin_amdgcn_s_getreg(XCC_ID_HWREG) & 15;
    if (xcc >= nx) return;
    const u32 tiles = 1u << (log_n - TL);
    const u32 r0 = (u32)((u64)nrefs * xcc / nx), mine = (u32)((u64)nrefs * (xcc + 1) / nx) - r0;
    u32* const done = ctl + FUSED_CTL_HEAD + r0;
    for (;;) {
        __syncthreads();  // the previous ticket's LDS use and s_ticket reads are over
        if (threadIdx.x == 0)
            s_ticket = __hip_atomic_fetch_add(ctl + xcc * CTL_STRIDE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const u32 t = s_ticket, s = t >> (log_n - TL), w = t & (tiles - 1);
        if (s >= mine + lag) return;
        // The thread index, opaque per ticket: otherwise the compiler keeps every address derived
        // from it for both passes live across the loop (about 290 VGPRs instead of the passes' own).
        u32 tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (s < mine) {
            const u32 ref_i = r0 + s;
            const LimbRef ref = refs[ref_i];
            const PrimeConst* pc = pcs + ref.prime;
            const u64 toff = (u64)ref.prime << log_n;
            if (pc->q < FP_PRIME_BOUND) {
                const FpAr ar{twf + toff, pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                              rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
                fwd_cols<MA>(ar, ref.src, ref.dst, lds, log_n, w, tid);
            } else {
                const IntAr ar{tw + toff, pc->q, pc->n_inv, pc->n_inv_s, pc->w_n_inv, pc->w_n_inv_s};
                fwd_cols<MA>(ar, ref.src, ref.dst, lds, log_n, w, tid);
            }
            __builtin_amdgcn_s_waitcnt(WAITCNT_VM0);  // this thread's stores have reached the L2
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_fetch_add(done + s, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (s >= lag && s - lag < mine) {
            const u32 m = s - lag, ref_i = r0 + m;
            if (threadIdx.x == 0) {
                u32 spins = 0;
                while (__hip_atomic_load(done + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tiles) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++spins > FUSED_SPIN_MAX) {
                        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __syncthreads();
            const LimbRef ref = refs[ref_i];
            const PrimeConst* pc = pcs + ref.prime;
            const u64 toff = (u64)ref.prime << log_n;
            if (pc->q < FP_PRIME_BOUND) {
                const FpAr ar{twf + toff, pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                              rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
                fwd_rows<MB, FpAr, true>(ar, ref.dst, ref.dst, lds, log_n, FIN ? fins + ref_i : nullptr, pc->q, w, tid);
            } else {
                const IntAr ar{tw + toff, pc->q, pc->n_inv, pc->n_inv_s, pc->w_n_inv, pc->w_n_inv_s};
                fwd_rows<MB, IntAr, true>(ar, ref.dst, ref.dst, lds, log_n, FIN ? fins + ref_i : nullptr, pc->q, w, tid);
            }
        }
    }
}

// The inverse transform in one launch, with the same ticket scheme: ticket t runs row tile w of the
// chunk's limb s (the first inverse pass, src -> dst), then column tile w of its limb s - lag
// (in place on dst), once every row tile of that limb has been counted done.
template <u32 MA, u32 MB>
__global__ __launch_bounds__(NT) void k_inv_fused(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                  const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                  const FTw* __restrict__ rt, u32 log_n, u32 nrefs, u32* ctl, u32 nx,
                                                  u32 lag, u32* err) {
    __shared__ u64 lds[LDS_WORDS];
    __shared__ u32 s_ticket;
    const u32 xcc = __builtin_amdgcn_s_getreg(XCC_ID_HWREG) & 15;
    if (xcc >= nx) return;
    const u32 tiles = 1u << (log_n - TL);
    const u32 r0 = (u32)((u64)nrefs * xcc / nx), mine = (u32)((u64)nrefs * (xcc + 1) / nx) - r0;
    u32* const done = ctl + FUSED_CTL_HEAD + r0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0)
            s_ticket = __hip_atomic_fetch_add(ctl + xcc * CTL_STRIDE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const u32 t = s_ticket, s = t >> (log_n - TL), w = t & (tiles - 1);
        if (s >= mine + lag) return;
        u32 tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if (s < mine) {
            const LimbRef ref = refs[r0 + s];
            const PrimeConst* pc = pcs + ref.prime;
            const u64 toff = (u64)ref.prime << log_n;
            if (pc->q < FP_PRIME_BOUND) {
                const FpAr ar{twf + toff, pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                              rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
                inv_rows<MB>(ar, ref.src, ref.dst, lds, log_n, w, tid);
            } else {
                const IntAr ar{tw + toff, pc->q, pc->n_inv, pc->n_inv_s, pc->w_n_inv, pc->w_n_inv_s};
                inv_rows<MB>(ar, ref.src, ref.dst, lds, log_n, w, tid);
            }
            __builtin_amdgcn_s_waitcnt(WAITCNT_VM0);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_fetch_add(done + s, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (s >= lag && s - lag < mine) {
            const u32 m = s - lag;
            if (threadIdx.x == 0) {
                u32 spins = 0;
                while (__hip_atomic_load(done + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < tiles) {
                    __builtin_amdgcn_s_sleep(2);
                    if (++spins > FUSED_SPIN_MAX) {
                        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                }
            }
            __syncthreads();
            const LimbRef ref = refs[r0 + m];
            const PrimeConst* pc = pcs + ref.prime;
            const u64 toff = (u64)ref.prime << log_n;
            if (pc->q < FP_PRIME_BOUND) {
                const FpAr ar{twf + toff, pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                              rt + ((u64)ref.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
                inv_cols<MA, FpAr, true>(ar, ref.dst, lds, log_n, w, tid);
            } else {
                const IntAr ar{tw + toff, pc->q, pc->n_inv, pc->n_inv_s, pc->w_n_inv, pc->w_n_inv_s};
                inv_cols<MA, IntAr, true>(ar, ref.dst, lds, log_n, w, tid);
            }
        }
    }
}

// XCD ids of a probe grid (one word per workgroup), for the check at context creation.
__global__ __launch_bounds__(64) void k_xcc_ids(u32* out) {
    if (threadIdx.x == 0) glb(out)[blockIdx.x] = __builtin_amdgcn_s_getreg(XCC_ID_HWREG) & 15;
}

// The fused row pass + inner product with mod_down's finish of both sums (RowIpFinJob); the same
// block order as k_rows_ip.
template <u32 M>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_rows_ip_fin(const RowIpFinJob* __restrict__ jobs,
                                                    const PrimeConst* __restrict__ pcs, const FTw* __restrict__ twf,
                                                    const FTw* __restrict__ rt, u32 log_n, u32 njobs) {
    __shared__ u64 lds[LDS_WORDS];
    const u32 tiles = 1u << (log_n - TL), ji = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    (void)njobs;
    const RowIpFinJob J = jobs[ji];
    const PrimeConst* pc = pcs + J.ip.prime;
    const FpAr ar{twf + ((u64)J.ip.prime << log_n), pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                  rt + ((u64)J.ip.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
    double sb[16], sa[16];
    rows_ip_sums<M>(ar, J.ip, lds, log_n, tile, sb, sa);
    finish_sum<M>(ar, sb, J.zb, J.out_b, J.add_b, J.inv, lds, log_n, tile);
    finish_sum<M>(ar, sa, J.za, J.out_a, J.add_a, J.inv, lds, log_n, tile);
}

template <u32 M>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3))) void k_rows_ip(const RowIpJob* __restrict__ jobs, const PrimeConst* __restrict__ pcs,
                                                const FTw* __restrict__ twf, const FTw* __restrict__ rt, u32 log_n,
                                                u32 njobs) {
    __shared__ u64 lds[LDS_WORDS];
    // Linear order: tile t of every job runs on XCD t % 8 (workgroups are dealt round-robin), and
    // jobs are ordered by limb, so the jobs reading one key tile share that XCD's L2.
    const u32 tiles = 1u << (log_n - TL), ji = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    (void)njobs;
    const RowIpJob J = jobs[ji];
    const PrimeConst* pc = pcs + J.prime;
    const FpAr ar{twf + ((u64)J.prime << log_n), pc->qf, pc->qinv, pc->n_inv_f, pc->w_n_inv_f,
                  rt + ((u64)J.prime << (log_n - log_n / 2)) * ROW_FACTOR_SLOTS};
    rows_ip<M>(ar, J, lds, log_n, tile);
}

}  // namespace fast

bool ntt_fast_supported(u32 log_n) { return log_n >= 12 && log_n <= 16; }

template <u32 M>
static void launch_one(bool fwd, bool cols, const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s,
                       const FinishJob* fins = nullptr, const YForm* yf = nullptr) {
    dim3 grid((1u << (t.log_n - fast::TL)) * nrefs);
    if (fwd && cols)
        hipLaunchKernelGGL(fast::k_fwd_cols<M>, grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf, t.log_n, nrefs);
    else if (fwd && fins)
        hipLaunchKernelGGL((fast::k_fwd_rows<M, true>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf,
                           t.log_n, nrefs, fins);
    else if (fwd)
        hipLaunchKernelGGL((fast::k_fwd_rows<M, false>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf,
                           t.log_n, nrefs, fins);
    else if (cols && yf)
        hipLaunchKernelGGL((fast::k_inv_cols<M, true>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.inv, t.inv_f, t.ri,
                           t.log_n, nrefs, yf);
    else if (cols)
        hipLaunchKernelGGL((fast::k_inv_cols<M, false>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.inv, t.inv_f, t.ri,
                           t.log_n, nrefs, yf);
    else
        hipLaunchKernelGGL(fast::k_inv_rows<M>, grid, dim3(fast::NT), 0, s, refs, t.pc, t.inv, t.inv_f, t.ri, t.log_n,
                           nrefs);
}

template <u32 MA, u32 MB>
static void launch_fused_ab(const NttTables& t, const LimbRef* refs, u32 nrefs, const FinishJob* fin, u32* ctl,
                            const FusedCfg& f, hipStream_t s) {
    if (fin)
        hipLaunchKernelGGL((fast::k_fwd_fused<MA, MB, true>), dim3(f.grid), dim3(fast::NT), 0, s, refs, t.pc, t.fwd,
                           t.fwd_f, t.rf, t.log_n, nrefs, fin, ctl, f.nx, f.lag, f.err);
    else
        hipLaunchKernelGGL((fast::k_fwd_fused<MA, MB, false>), dim3(f.grid), dim3(fast::NT), 0, s, refs, t.pc, t.fwd,
                           t.fwd_f, t.rf, t.log_n, nrefs, fin, ctl, f.nx, f.lag, f.err);
}

void ntt_fast_forward_fused(const NttTables& t, const LimbRef* refs, u32 nrefs, const FinishJob* fin, u32* ctl,
                            const FusedCfg& f, hipStream_t s) {
    if (nrefs == 0) return;
    const u32 B = t.log_n / 2, A = t.log_n - B;
    if (A == 6 && B == 6) launch_fused_ab<6, 6>(t, refs, nrefs, fin, ctl, f, s);
    else if (A == 7 && B == 6) launch_fused_ab<7, 6>(t, refs, nrefs, fin, ctl, f, s);
    else if (A == 7 && B == 7) launch_fused_ab<7, 7>(t, refs, nrefs, fin, ctl, f, s);
    else if (A == 8 && B == 7) launch_fused_ab<8, 7>(t, refs, nrefs, fin, ctl, f, s);
    else launch_fused_ab<8, 8>(t, refs, nrefs, fin, ctl, f, s);
}

template <u32 MA, u32 MB>
static void launch_inv_fused_ab(const NttTables& t, const LimbRef* refs, u32 nrefs, u32* ctl, const FusedCfg& f,
                                hipStream_t s) {
    hipLaunchKernelGGL((fast::k_inv_fused<MA, MB>), dim3(f.grid), dim3(fast::NT), 0, s, refs, t.pc, t.inv, t.inv_f, t.ri,
                       t.log_n, nrefs, ctl, f.nx, f.lag, f.err);
}

void ntt_fast_inverse_fused(const NttTables& t, const LimbRef* refs, u32 nrefs, u32* ctl, const FusedCfg& f,
                            hipStream_t s) {
    if (nrefs == 0) return;
    const u32 B = t.log_n / 2, A = t.log_n - B;
    if (A == 6 && B == 6) launch_inv_fused_ab<6, 6>(t, refs, nrefs, ctl, f, s);
    else if (A == 7 && B == 6) launch_inv_fused_ab<7, 6>(t, refs, nrefs, ctl, f, s);
    else if (A == 7 && B == 7) launch_inv_fused_ab<7, 7>(t, refs, nrefs, ctl, f, s);
    else if (A == 8 && B == 7) launch_inv_fused_ab<8, 7>(t, refs, nrefs, ctl, f, s);
    else launch_inv_fused_ab<8, 8>(t, refs, nrefs, ctl, f, s);
}

u32 fused_occupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fast::k_fwd_fused<8, 8, false>, fast::NT, 0) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return (u32)blocks;
}

u32 xcc_layout_check(hipStream_t s) {
    constexpr u32 G = 4096;
    u32* d = nullptr;
    if (hipMalloc(&d, G * sizeof(u32)) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    std::vector<u32> h(G, 0xffffffffu);
    hipLaunchKernelGGL(fast::k_xcc_ids, dim3(G), dim3(64), 0, s, d);
    const bool ok = hipMemcpyAsync(h.data(), d, G * sizeof(u32), hipMemcpyDeviceToHost, s) == hipSuccess &&
                    hipStreamSynchronize(s) == hipSuccess;
    (void)hipFree(d);
    if (!ok) {
        (void)hipGetLastError();
        return 0;
    }
    // Ids 0..nx-1 each present (the round-robin dispatch may start at any XCD, so the id of a
    // workgroup is not its index modulo nx; the one-launch transform reads the id itself).
    u32 nx = 0;
    u32 seen[16] = {0};
    for (u32 v : h) {
        if (v >= 8) return 0;
        nx = std::max(nx, v + 1);
        ++seen[v];
    }
    for (u32 x = 0; x < nx; ++x)
        if (seen[x] < G / (4 * nx)) return 0;
    return nx;
}

static void launch_m(u32 m, bool fwd, bool cols, const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s,
                     const FinishJob* fins = nullptr, const YForm* yf = nullptr) {
    switch (m) {
        case 6: launch_one<6>(fwd, cols, t, refs, nrefs, s, fins, yf); break;
        case 7: launch_one<7>(fwd, cols, t, refs, nrefs, s, fins, yf); break;
        default: launch_one<8>(fwd, cols, t, refs, nrefs, s, fins, yf); break;
    }
}

// Split N = 2^(A+B): B = floor(log N / 2) row stages, A column stages (6..8 each).
void ntt_fast_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s) {
    const u32 B = t.log_n / 2, A = t.log_n - B;
    launch_m(A, true, true, t, refs, nrefs, s);
    launch_m(B, true, false, t, refs, nrefs, s);
}

void ntt_fast_forward_finish(const NttTables& t, const LimbRef* refs, const FinishJob* fin, u32 nrefs,
                             hipStream_t s) {
    const u32 B = t.log_n / 2, A = t.log_n - B;
    launch_m(A, true, true, t, refs, nrefs, s);
    launch_m(B, true, false, t, refs, nrefs, s, fin);
}

void ntt_fast_forward_cols(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s) {
    launch_m(t.log_n - t.log_n / 2, true, true, t, refs, nrefs, s);
}

void ntt_fast_forward_rows(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s) {
    launch_m(t.log_n / 2, true, false, t, refs, nrefs, s);
}

void rows_ip_fin_launch(const NttTables& t, const RowIpFinJob* jobs, u32 njobs, hipStream_t s) {
    if (njobs == 0) return;
    dim3 grid((1u << (t.log_n - fast::TL)) * njobs);
    switch (t.log_n / 2) {
        case 6: hipLaunchKernelGGL(fast::k_rows_ip_fin<6>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        case 7: hipLaunchKernelGGL(fast::k_rows_ip_fin<7>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        default: hipLaunchKernelGGL(fast::k_rows_ip_fin<8>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
    }
}

void rows_ip_launch(const NttTables& t, const RowIpJob* jobs, u32 njobs, hipStream_t s) {
    if (njobs == 0) return;
    dim3 grid((1u << (t.log_n - fast::TL)) * njobs);
    switch (t.log_n / 2) {
        case 6: hipLaunchKernelGGL(fast::k_rows_ip<6>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        case 7: hipLaunchKernelGGL(fast::k_rows_ip<7>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        default: hipLaunchKernelGGL(fast::k_rows_ip<8>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
    }
}

void ntt_fast_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s, const YForm* yf) {
    const u32 B = t.log_n / 2, A = t.log_n - B;
    launch_m(B, false, false, t, refs, nrefs, s);
    launch_m(A, false, true, t, refs, nrefs, s, nullptr, yf);
}

}  // namespace ckh
