// Negacyclic NTT / inverse NTT for gfx950, batched over limbs.
//
// Conventions (crates/ckks-train-core/src/ntt.rs:4-26): forward is Cooley-Tukey from natural
// order to bit-reversed order with roots[bitrev(i)] = psi^i; inverse is Gentleman-Sande from
// bit-reversed to natural order with the N^-1 scaling fused into the last stage. Both take
// and return fully reduced residues, so every correct schedule returns the same words.
//
// Schedule: a limb of N = 2^(A+B) words is viewed as 2^A rows of 2^B columns. The first A
// forward stages (strides >= 2^B) only mix elements of one column; the last B stages only mix
// elements of one row. One pass handles a 4096-word tile in LDS: the column pass a tile of
// 4096 / 2^A adjacent columns, the row pass a tile of 4096 / 2^B rows. The inverse runs rows
// first, then columns. For N <= 4096 one row pass does everything (A = 0). Every butterfly's
// twiddle is roots[2^s + (i >> (logN - s))] for stage s and element index i, read from a
// per-prime table of (w, w') pairs. Lazy ranges follow Harvey: forward values stay in [0, 4q),
// inverse values in [0, 2q).
#include "common.h"
#include "ntt.h"

namespace ckh {

constexpr u32 TILE_LOG = 12;
constexpr u32 TILE = 1u << TILE_LOG;
constexpr u32 NTHREADS = 256;

static inline void split_log(u32 log_n, u32& A, u32& B) {
    if (log_n <= TILE_LOG) {
        A = 0;
        B = log_n;
    } else {
        B = log_n / 2;
        A = log_n - B;
    }
}

// Forward CT butterfly with lazy reduction: x, y in [0, 4q) -> [0, 4q).
CKH_DEV void ct_bfly(u64& x, u64& y, Tw w, u64 q) {
    const u64 two_q = 2 * q;
    u64 xr = x >= two_q ? x - two_q : x;
    u64 t = shoup_lazy(y, w.w, w.ws, q);
    x = xr + t;
    y = xr + two_q - t;
}

CKH_DEV u64 reduce4q(u64 r, u64 q) {
    r = r >= 2 * q ? r - 2 * q : r;
    return r >= q ? r - q : r;
}

// Inverse GS butterfly: x, y in [0, 2q) -> [0, 2q).
CKH_DEV void gs_bfly(u64& x, u64& y, Tw w, u64 q) {
    const u64 two_q = 2 * q;
    u64 s = x + y;
    u64 d = x + two_q - y;
    x = s >= two_q ? s - two_q : s;
    y = shoup_lazy(d, w.w, w.ws, q);
}

// Last inverse stage with N^-1 fused: outputs in [0, q).
CKH_DEV void gs_last(u64& x, u64& y, const PrimeConst& c) {
    const u64 q = c.q;
    u64 s = x + y;
    u64 d = x + 2 * q - y;
    x = csub(shoup_lazy(s, c.n_inv, c.n_inv_s, q), q);
    y = csub(shoup_lazy(d, c.w_n_inv, c.w_n_inv_s, q), q);
}

// ---------------------------------------------------------------- forward

__global__ __launch_bounds__(NTHREADS) void k_ntt_fwd_cols(const LimbRef* __restrict__ refs,
                                                           const PrimeConst* __restrict__ pcs,
                                                           const Tw* __restrict__ tw, u32 log_n, u32 A) {
    __shared__ u64 s[TILE];
    const u32 B = log_n - A;
    const LimbRef ref = refs[blockIdx.y];
    const u64 q = pcs[ref.prime].q;
    const Tw* roots = tw + ((u64)ref.prime << log_n);
    const u32 tc_log = TILE_LOG - A;
    const u32 TC = 1u << tc_log;
    const u32 c0 = blockIdx.x * TC;
    const auto in = glb(ref.src);
    const auto out = glb(ref.dst);
    for (u32 e = threadIdx.x; e < TILE; e += NTHREADS) {
        const u32 r = e >> tc_log, cc = e & (TC - 1);
        s[e] = in[((u64)r << B) + c0 + cc];
    }
    __syncthreads();
    for (u32 st = 0; st < A; ++st) {
        const u32 hl = A - 1 - st;  // log2 of the row distance
        for (u32 b = threadIdx.x; b < TILE / 2; b += NTHREADS) {
            const u32 cc = b & (TC - 1), p = b >> tc_log;
            const u32 r = ((p >> hl) << (hl + 1)) | (p & ((1u << hl) - 1));
            const Tw w = roots[(1u << st) + (r >> (A - st))];
            u64 x = s[(r << tc_log) + cc], y = s[((r + (1u << hl)) << tc_log) + cc];
            ct_bfly(x, y, w, q);
            s[(r << tc_log) + cc] = x;
            s[((r + (1u << hl)) << tc_log) + cc] = y;
        }
        __syncthreads();
    }
    for (u32 e = threadIdx.x; e < TILE; e += NTHREADS) {
        const u32 r = e >> tc_log, cc = e & (TC - 1);
        out[((u64)r << B) + c0 + cc] = s[e];
    }
}

__global__ __launch_bounds__(NTHREADS) void k_ntt_fwd_rows(const LimbRef* __restrict__ refs,
                                                           const PrimeConst* __restrict__ pcs,
                                                           const Tw* __restrict__ tw, u32 log_n, u32 A,
                                                           int src_is_dst) {
    __shared__ u64 s[TILE];
    const u32 B = log_n - A;
    const LimbRef ref = refs[blockIdx.y];
    const u64 q = pcs[ref.prime].q;
    const Tw* roots = tw + ((u64)ref.prime << log_n);
    const u32 tr_log = (TILE_LOG - B) < A ? (TILE_LOG - B) : A;  // rows per tile (log)
    const u32 elems = 1u << (tr_log + B);
    const u64 base = (u64)blockIdx.x << (tr_log + B);  // first element of the tile
    const auto in = glb((src_is_dst ? ref.dst : ref.src) + base);
    const auto out = glb(ref.dst + base);
    for (u32 e = threadIdx.x; e < elems; e += NTHREADS) s[e] = in[e];
    __syncthreads();
    for (u32 u = 0; u < B; ++u) {
        const u32 st = A + u;
        const u32 hl = B - 1 - u;  // log2 of the column distance
        for (u32 b = threadIdx.x; b < elems / 2; b += NTHREADS) {
            const u32 rr = b >> (B - 1), p = b & ((1u << (B - 1)) - 1);
            const u32 c = ((p >> hl) << (hl + 1)) | (p & ((1u << hl) - 1));
            const u32 li = (rr << B) + c;
            const u64 gi = base + li;
            const Tw w = roots[(1u << st) + (u32)(gi >> (B - u))];
            u64 x = s[li], y = s[li + (1u << hl)];
            ct_bfly(x, y, w, q);
            if (u == B - 1) {
                x = reduce4q(x, q);
                y = reduce4q(y, q);
            }
            s[li] = x;
            s[li + (1u << hl)] = y;
        }
        __syncthreads();
    }
    for (u32 e = threadIdx.x; e < elems; e += NTHREADS) out[e] = s[e];
}

// ---------------------------------------------------------------- inverse

__global__ __launch_bounds__(NTHREADS) void k_ntt_inv_rows(const LimbRef* __restrict__ refs,
                                                           const PrimeConst* __restrict__ pcs,
                                                           const Tw* __restrict__ tw, u32 log_n, u32 A) {
    __shared__ u64 s[TILE];
    const u32 B = log_n - A;
    const LimbRef ref = refs[blockIdx.y];
    const PrimeConst pc = pcs[ref.prime];
    const u64 q = pc.q;
    const Tw* iroots = tw + ((u64)ref.prime << log_n);
    const u32 tr_log = (TILE_LOG - B) < A ? (TILE_LOG - B) : A;
    const u32 elems = 1u << (tr_log + B);
    const u64 base = (u64)blockIdx.x << (tr_log + B);
    const auto in = glb(ref.src + base);
    const auto out = glb(ref.dst + base);
    for (u32 e = threadIdx.x; e < elems; e += NTHREADS) s[e] = in[e];
    __syncthreads();
    for (u32 u = 0; u < B; ++u) {
        const u32 hl = u;  // pair distance 2^u
        const bool last = (u == log_n - 1);
        for (u32 b = threadIdx.x; b < elems / 2; b += NTHREADS) {
            const u32 rr = b >> (B - 1), p = b & ((1u << (B - 1)) - 1);
            const u32 c = ((p >> hl) << (hl + 1)) | (p & ((1u << hl) - 1));
            const u32 li = (rr << B) + c;
            u64 x = s[li], y = s[li + (1u << hl)];
            if (last) {
                gs_last(x, y, pc);
            } else {
                const u64 gi = base + li;
                const Tw w = iroots[(1u << (log_n - 1 - u)) + (u32)(gi >> (u + 1))];
                gs_bfly(x, y, w, q);
            }
            s[li] = x;
            s[li + (1u << hl)] = y;
        }
        __syncthreads();
    }
    for (u32 e = threadIdx.x; e < elems; e += NTHREADS) out[e] = s[e];
}

__global__ __launch_bounds__(NTHREADS) void k_ntt_inv_cols(const LimbRef* __restrict__ refs,
                                                           const PrimeConst* __restrict__ pcs,
                                                           const Tw* __restrict__ tw, u32 log_n, u32 A) {
    __shared__ u64 s[TILE];
    const u32 B = log_n - A;
    const LimbRef ref = refs[blockIdx.y];
    const PrimeConst pc = pcs[ref.prime];
    const u64 q = pc.q;
    const Tw* iroots = tw + ((u64)ref.prime << log_n);
    const u32 tc_log = TILE_LOG - A;
    const u32 TC = 1u << tc_log;
    const u32 c0 = blockIdx.x * TC;
    const auto io = glb(ref.dst);
    for (u32 e = threadIdx.x; e < TILE; e += NTHREADS) {
        const u32 r = e >> tc_log, cc = e & (TC - 1);
        s[e] = io[((u64)r << B) + c0 + cc];
    }
    __syncthreads();
    for (u32 v = 0; v < A; ++v) {
        const u32 u = B + v;
        const bool last = (v == A - 1);
        for (u32 b = threadIdx.x; b < TILE / 2; b += NTHREADS) {
            const u32 cc = b & (TC - 1), p = b >> tc_log;
            const u32 r = ((p >> v) << (v + 1)) | (p & ((1u << v) - 1));
            u64 x = s[(r << tc_log) + cc], y = s[((r + (1u << v)) << tc_log) + cc];
            if (last) {
                gs_last(x, y, pc);
            } else {
                const Tw w = iroots[(1u << (log_n - 1 - u)) + (r >> (v + 1))];
                gs_bfly(x, y, w, q);
            }
            s[(r << tc_log) + cc] = x;
            s[((r + (1u << v)) << tc_log) + cc] = y;
        }
        __syncthreads();
    }
    for (u32 e = threadIdx.x; e < TILE; e += NTHREADS) {
        const u32 r = e >> tc_log, cc = e & (TC - 1);
        io[((u64)r << B) + c0 + cc] = s[e];
    }
}

// ---------------------------------------------------------------- launchers

void ntt_forward(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream) {
    if (nrefs == 0) return;
    if (ntt_fast_supported(t.log_n)) return ntt_fast_forward(t, refs, nrefs, stream);
    u32 A, B;
    split_log(t.log_n, A, B);
    if (A > 0) {
        dim3 grid(1u << (t.log_n - TILE_LOG), nrefs);
        hipLaunchKernelGGL(k_ntt_fwd_cols, grid, dim3(NTHREADS), 0, stream, refs, t.pc, t.fwd, t.log_n, A);
        hipLaunchKernelGGL(k_ntt_fwd_rows, grid, dim3(NTHREADS), 0, stream, refs, t.pc, t.fwd, t.log_n, A, 1);
    } else {
        hipLaunchKernelGGL(k_ntt_fwd_rows, dim3(1, nrefs), dim3(NTHREADS), 0, stream, refs, t.pc, t.fwd, t.log_n,
                           0u, 0);
    }
}

void ntt_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t stream) {
    if (nrefs == 0) return;
    if (ntt_fast_supported(t.log_n)) return ntt_fast_inverse(t, refs, nrefs, stream);
    u32 A, B;
    split_log(t.log_n, A, B);
    if (A > 0) {
        dim3 grid(1u << (t.log_n - TILE_LOG), nrefs);
        hipLaunchKernelGGL(k_ntt_inv_rows, grid, dim3(NTHREADS), 0, stream, refs, t.pc, t.inv, t.log_n, A);
        hipLaunchKernelGGL(k_ntt_inv_cols, grid, dim3(NTHREADS), 0, stream, refs, t.pc, t.inv, t.log_n, A);
    } else {
        hipLaunchKernelGGL(k_ntt_inv_rows, dim3(1, nrefs), dim3(NTHREADS), 0, stream, refs, t.pc, t.inv, t.log_n,
                           0u);
    }
}

}  // namespace ckh
