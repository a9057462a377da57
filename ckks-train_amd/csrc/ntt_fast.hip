// Register-radix NTT passes for 2^12 <= N <= 2^16 (gfx950).
//
// Same transform as ntt.hip (crates/ckks-train-core/src/ntt.rs:141-232; outputs are unique
// residues), scheduled for CDNA: a limb of N = 2^(A+B) words is 2^A rows x 2^B columns; the
// column pass runs the first A forward stages, the row pass the last B. A pass works on a tile
// of 4096 words with 256 threads holding 16 words each in registers: four stages run in
// registers (radix 16), one LDS exchange regroups the words, and the remaining stages of the
// pass run in registers again, so a pass has one barrier for its exchange instead of one per
// stage. The row pass stages its output through LDS so every global store is coalesced; the
// inverse row pass stages its input likewise.
//
// Arithmetic, chosen per limb (uniform per workgroup):
//  * q < 2^50 (FpAr): residues as f64 integers. Twiddle product t = y w mod q via
//    h = RN(y w), l = fma(y, w, -h) (exact: y w = h + l), k = rint(y RN(w/q)),
//    t = fma(-k, q, h) + l. With |y| <= 2^51 the quotient error is below 1, |t| <= q, and
//    h - k q, t are integers below 2^53, so every step is exact. Forward: x is reduced to
//    |x| <= q/2 by x - rint(x/q) q before x +- t, keeping every value within 1.5 q < 2^51.
//    Inverse: |x|, |y| <= q throughout. Between the two passes a limb holds these signed
//    values as f64 bits; the last stage writes canonical residues.
//  * q >= 2^50 (IntAr): Harvey butterflies with 64-bit Shoup products, values in [0, 4q)
//    (forward) or [0, 2q) (inverse), as in ntt.hip.
#include <algorithm>
#include <vector>

#include "basis.h"
#include "common.h"
#include "ntt.h"

namespace ckh {
namespace fast {

constexpr u32 TL = 12;
constexpr u32 TILE = 1u << TL;
constexpr u32 NT = 256;
constexpr u32 LDS_WORDS = TILE + TILE / 16;
// s_getreg operand for HW_REG_XCC_ID (register 20), bits 0..3: the XCD the wave runs on.
constexpr int XCC_ID_HWREG = 20 | (0 << 6) | ((4 - 1) << 11);
// s_waitcnt operand: vmcnt(0), expcnt and lgkmcnt not waited for.
constexpr int WAITCNT_VM0 = 0x0F70;
constexpr u32 FUSED_SPIN_MAX = 1u << 22;

// Padded LDS index: one spare word per 16 spreads strided accesses over the banks.
CKH_DEV u32 pad(u32 a) { return a + (a >> 4); }

// ---------------------------------------------------------------- f64 arithmetic (q < 2^50)

struct FpAr {
    typedef double V;
    typedef FTw W;
    const FTw* tab;
    double q, qinv;
    FTw ninv, wninv;
    const FTw* rt;  // this prime's row factors, [row][ROW_FACTOR_SLOTS]
    CKH_DEV V from_canon(u64 x) const { return u2d(x); }
    CKH_DEV V from_mid(u64 x) const { return __longlong_as_double((long long)x); }
    CKH_DEV u64 to_mid(V v) const { return (u64)__double_as_longlong(v); }
    CKH_DEV W tw(u32 idx) const { return tab[idx]; }
    // Row-pass twiddles tab[2^(A+s) + row 2^s + j0 + m], m < 2^D (j0 + 2^D <= 2^s): the table
    // holds psi^bitrev(i) and bit reversal maps disjoint bits to disjoint bits, so the entry is
    // the row factor rt[row][s] = tab[2^(A+s) + row 2^s] (one 128-byte line per row) times
    // tab[j0 + m] (among the first 2^M entries, cache resident). m = 0 with j0 = 0 is the
    // factor itself. Products are signed, of magnitude below 0.75 q, which every butterfly
    // accepts as a twiddle (w / q rounds twice: |k error| <= 3 |y| / (16 q) + 1/2 stays below 1
    // for |y| <= 2^51).
    template <u32 D>
    CKH_DEV void row_group(u32 /*full_base*/, u32 row, u32 s, u32 j0, W* g) const {
        const FTw r = rt[row * ROW_FACTOR_SLOTS + s];
#pragma unroll
        for (u32 m = 0; m < (1u << D); ++m) {
            if (m == 0 && j0 == 0) {
                g[0] = r;
                continue;
            }
            const FTw c = tab[j0 + m];
            const double w = fmulmod(r.w, c.w, c.wq, q);
            g[m] = {w, w * qinv};
        }
    }
    CKH_DEV void ct(V& x, V& y, const W& w) const {
        const double xr = freduce(x, q, qinv);
        const double t = fmulmod(y, w.w, w.wq, q);
        x = xr + t;
        y = xr - t;
    }
    // Without the reduction of x. Stages alternate (lazy on even stages of each pass): with
    // |t| <= (1/2 + |y| / (4 q)) q the magnitudes stay below 1.7 q after a reduced stage,
    // 2.6 q after a lazy one, and 3.7 q where two lazy stages meet at the pass boundary; every
    // value and every fmulmod intermediate stays an exact integer below 2^53.
    CKH_DEV void ct_lazy(V& x, V& y, const W& w) const {
        const double t = fmulmod(y, w.w, w.wq, q);
        const double x0 = x;
        x = x0 + t;
        y = x0 - t;
    }
    CKH_DEV u64 fwd_final(V v) const { return d2u(fcanon(v, q, qinv)); }
    CKH_DEV u64 fwd_final_f(V v) const { return (u64)__double_as_longlong(fcanon(v, q, qinv)); }
    CKH_DEV void gs(V& x, V& y, const W& w) const {
        const double s = x + y, d = x - y;
        x = freduce(s, q, qinv);
        y = fmulmod(d, w.w, w.wq, q);
    }
    CKH_DEV void gs_final(V& x, V& y) const {
        const double s = x + y, d = x - y;
        x = fcanon(fmulmod(s, ninv.w, ninv.wq, q), q, qinv);
        y = fcanon(fmulmod(d, wninv.w, wninv.wq, q), q, qinv);
    }
    CKH_DEV u64 canon_out(V v) const { return d2u(v); }
};

// ---------------------------------------------------------------- integer arithmetic

struct IntAr {
    typedef u64 V;
    typedef Tw W;
    const Tw* tab;
    u64 q;
    u64 n_inv, n_inv_s, w_n_inv, w_n_inv_s;
    CKH_DEV V from_canon(u64 x) const { return x; }
    CKH_DEV V from_mid(u64 x) const { return x; }
    CKH_DEV u64 to_mid(V v) const { return v; }
    CKH_DEV W tw(u32 idx) const { return tab[idx]; }
    template <u32 D>
    CKH_DEV void row_group(u32 full_base, u32, u32, u32, W* g) const {
#pragma unroll
        for (u32 m = 0; m < (1u << D); ++m) g[m] = tab[full_base + m];
    }
    CKH_DEV void ct(V& x, V& y, const W& w) const {
        const u64 two_q = 2 * q;
        const u64 xr = x >= two_q ? x - two_q : x;
        const u64 t = shoup_lazy(y, w.w, w.ws, q);
        x = xr + t;
        y = xr + two_q - t;
    }
    CKH_DEV void ct_lazy(V& x, V& y, const W& w) const { ct(x, y, w); }
    CKH_DEV u64 fwd_final_f(V v) const { return fwd_final(v); }  // not used (f64 primes only)
    CKH_DEV u64 fwd_final(V v) const {
        v = v >= 2 * q ? v - 2 * q : v;
        return v >= q ? v - q : v;
    }
    CKH_DEV void gs(V& x, V& y, const W& w) const {
        const u64 two_q = 2 * q;
        const u64 s = x + y, d = x + two_q - y;
        x = s >= two_q ? s - two_q : s;
        y = shoup_lazy(d, w.w, w.ws, q);
    }
    CKH_DEV void gs_final(V& x, V& y) const {
        const u64 s = x + y, d = x + 2 * q - y;
        x = csub(shoup_lazy(s, n_inv, n_inv_s, q), q);
        y = csub(shoup_lazy(d, w_n_inv, w_n_inv_s, q), q);
    }
    CKH_DEV u64 canon_out(V v) const { return v; }
};

// ---------------------------------------------------------------- row-pass twiddle groups
//
// In the row passes every thread needs, per stage, 2^D twiddles at consecutive table indices
// starting at a multiple of 2^D; Ar::group builds them (for the f64 path from one load).

// Forward butterfly of stage s of a pass (s is a constant after unrolling).
template <class Ar>
CKH_DEV void ct_stage(const Ar& ar, u32 s, typename Ar::V& x, typename Ar::V& y, const typename Ar::W& w) {
    if (s & 1)
        ar.ct(x, y, w);
    else
        ar.ct_lazy(x, y, w);
}

// Stage twiddles of the row passes: tab[full_base + m], m < 2^D, where full_base =
// 2^(A+s) + row 2^s + j0. D is a compile-time constant once the callers' loops are unrolled.
template <class Ar>
CKH_DEV void row_tw(const Ar& ar, u32 A, u32 row, u32 s, u32 j0, u32 D, typename Ar::W* g) {
    const u32 full = (1u << (A + s)) + (row << s) + j0;
    switch (D) {
        case 0: ar.template row_group<0>(full, row, s, j0, g); break;
        case 1: ar.template row_group<1>(full, row, s, j0, g); break;
        case 2: ar.template row_group<2>(full, row, s, j0, g); break;
        default: ar.template row_group<3>(full, row, s, j0, g); break;
    }
}

template <class Ar>
CKH_DEV void rows_round1_fwd(const Ar& ar, typename Ar::V (&v)[16], u32 A, u32 row) {
#pragma unroll
    for (u32 s = 0; s < 4; ++s) {
        typename Ar::W g[8];
        row_tw(ar, A, row, s, 0, s, g);
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            const u32 d = 1u << (3 - s);
            if (k & d) continue;
            ct_stage(ar, s, v[k], v[k + d], g[k >> (4 - s)]);
        }
    }
}

template <u32 M, class Ar>
CKH_DEV void rows_round2_fwd(const Ar& ar, typename Ar::V (&v)[16], u32 A, u32 row, u32 o) {
#pragma unroll
    for (u32 s = 4; s < M; ++s) {
        const u32 D = s + 4 - M, p = M - 1 - s;
        typename Ar::W g[8];
        row_tw(ar, A, row, s, o << D, D, g);
#pragma unroll
        for (u32 e = 0; e < 16; ++e) {
            if (e & (1u << p)) continue;
            ct_stage(ar, s, v[e], v[e + (1u << p)], g[e >> (p + 1)]);
        }
    }
}

// Inverse stage u of the row pass is row stage s = M - 1 - u (table index 2^(A+s) + row 2^s + j).
template <u32 M, class Ar>
CKH_DEV void rows_round1_inv(const Ar& ar, typename Ar::V (&v)[16], u32 A, u32 row, u32 o) {
    constexpr u32 R2 = M - 4;
#pragma unroll
    for (u32 u = 0; u < R2; ++u) {
        const u32 D = 3 - u;
        typename Ar::W g[8];
        row_tw(ar, A, row, M - 1 - u, o << D, D, g);
#pragma unroll
        for (u32 e = 0; e < 16; ++e) {
            if (e & (1u << u)) continue;
            ar.gs(v[e], v[e + (1u << u)], g[e >> (u + 1)]);
        }
    }
}

template <u32 M, class Ar>
CKH_DEV void rows_round2_inv(const Ar& ar, typename Ar::V (&v)[16], u32 A, u32 row) {
    constexpr u32 R2 = M - 4;
#pragma unroll
    for (u32 u = R2; u < M; ++u) {
        const u32 D = M - 1 - u, d = 1u << (u - R2);
        typename Ar::W g[8];
        row_tw(ar, A, row, M - 1 - u, 0, D, g);
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            if (k & d) continue;
            ar.gs(v[k], v[k + d], g[k >> (u + 1 - R2)]);
        }
    }
}

// ---------------------------------------------------------------- passes
//
// Thread layout (M = log2 of the sub-transform, 6..8; R2 = M - 4 bits handled in round 2;
// U = 4 - R2). "k-layout": a thread holds sub-transform indices o + (k << R2), k < 16 (the
// four top bits vary). "e-layout": it holds ((hh << U | u) << R2) | j, e = (u << R2) | j (the
// R2 low bits vary, plus U of the top bits).

template <u32 M, class Ar>
CKH_DEV void fwd_cols(const Ar& ar, const u64* __restrict__ in_p, u64* __restrict__ out_p, u64* lds, u32 log_n,
                      u32 tile, u32 tid = threadIdx.x) {
    const auto in = glb(in_p);
    const auto out = glb(out_p);
    constexpr u32 R2 = M - 4, U = 4 - R2, TCL = TL - M, TC = 1u << TCL, JM = (1u << R2) - 1;
    const u32 B = log_n - M;
    const u32 cc = tid & (TC - 1), o = tid >> TCL;
    const u32 c0 = tile * TC;
    typename Ar::V v[16];
#pragma unroll
    for (u32 k = 0; k < 16; ++k) v[k] = ar.from_canon(in[((u64)(o + (k << R2)) << B) + c0 + cc]);
#pragma unroll
    for (u32 s = 0; s < 4; ++s) {
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            const u32 d = 1u << (3 - s);
            if (k & d) continue;
            ct_stage(ar, s, v[k], v[k + d], ar.tw((1u << s) + (k >> (4 - s))));
        }
    }
#pragma unroll
    for (u32 k = 0; k < 16; ++k) lds[pad(((o + (k << R2)) << TCL) + cc)] = ar.to_mid(v[k]);
    __syncthreads();
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        v[e] = ar.from_mid(lds[pad((r << TCL) + cc)]);
    }
#pragma unroll
    for (u32 s = 4; s < M; ++s) {
        const u32 p = M - 1 - s;
#pragma unroll
        for (u32 e = 0; e < 16; ++e) {
            if (e & (1u << p)) continue;
            const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
            ct_stage(ar, s, v[e], v[e + (1u << p)], ar.tw((1u << s) + (r >> (M - s))));
        }
    }
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        out[((u64)r << B) + c0 + cc] = ar.to_mid(v[e]);
    }
}

// The row pass of one tile, its canonical outputs left in LDS in natural order (word i of the
// tile at pad(i)): as u64 residues, or (F64) as the f64 bits of the residue.
// A word of a column-pass output written by another workgroup of the same launch: loaded at agent
// scope (the sc1 bit), so the load misses this CU's vector L1 (which may hold an older copy of the
// line) and reads the XCD's L2, where the writer's completed stores are.
CKH_DEV u64 load_coherent(const u64* p) {
    return __hip_atomic_load(glb(const_cast<u64*>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <u32 M, bool F64, class Ar, bool COH = false>
CKH_DEV void fwd_rows_to_lds(const Ar& ar, const u64* __restrict__ in_p, u64* lds, u32 log_n, u32 tile,
                             u32 tid = threadIdx.x) {
    const auto in = glb(in_p);
    constexpr u32 R2 = M - 4, U = 4 - R2, TRL = TL - M, JM = (1u << R2) - 1;
    const u32 A = log_n - M;
    const u32 o = tid & JM, rr = tid >> R2;
    const u32 row = (tile << TRL) + rr;
    const u64 base = (u64)row << M;
    typename Ar::V v[16];
#pragma unroll
    for (u32 k = 0; k < 16; ++k)
        v[k] = ar.from_mid(COH ? load_coherent(in_p + base + o + (k << R2)) : in[base + o + (k << R2)]);
    // Round 1: stage A + s uses tab[2^(A+s) + row 2^s + m], m = k >> (4 - s) < 2^s.
    rows_round1_fwd<Ar>(ar, v, A, row);
#pragma unroll
    for (u32 k = 0; k < 16; ++k) lds[pad((rr << M) + o + (k << R2))] = ar.to_mid(v[k]);
    __syncthreads();
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 c = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        v[e] = ar.from_mid(lds[pad((rr << M) + c)]);
    }
    // Round 2: the thread's column c = 16 o + e, so stage A + s uses tab[2^(A+s) + row 2^s +
    // o 2^D + m] with D = s + 4 - M and m = e >> (M - s) < 2^D.
    rows_round2_fwd<M, Ar>(ar, v, A, row, o);
    __syncthreads();
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 c = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        lds[pad((rr << M) + c)] = F64 ? ar.fwd_final_f(v[e]) : ar.fwd_final(v[e]);
    }
    __syncthreads();
}

template <u32 M, class Ar, bool COH = false>
CKH_DEV void fwd_rows(const Ar& ar, const u64* __restrict__ in, u64* __restrict__ out_p, u64* lds, u32 log_n,
                      const FinishJob* fin, u64 q, u32 tile, u32 tid = threadIdx.x) {
    const auto out = glb(out_p);
    const u64 tbase = (u64)tile << TL;
    if (fin) {
        // finish of rescale / mod_down: (a - t) * inv (+ add); the a (and add) words are loaded
        // before the transform so their latency overlaps it
        const auto fa = glb(fin->a), fadd = glb(fin->add);
        const auto fout = glb(fin->out);
        const bool has_add = fin->add != nullptr;
        const u64 inv = fin->inv, inv_s = fin->inv_s;
        u64 av[16], dv[16];
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            av[k] = fa[tbase + tid + k * NT];
            dv[k] = has_add ? fadd[tbase + tid + k * NT] : 0;
        }
        fwd_rows_to_lds<M, false, Ar, COH>(ar, in, lds, log_n, tile, tid);
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            u64 v = shoup(sub_mod(av[k], lds[pad(tid + k * NT)], q), inv, inv_s, q);
            if (has_add) v = add_mod(v, dv[k], q);
            fout[tbase + tid + k * NT] = v;
        }
        return;
    }
    fwd_rows_to_lds<M, false, Ar, COH>(ar, in, lds, log_n, tile, tid);
#pragma unroll
    for (u32 k = 0; k < 16; ++k) {
        const u32 idx = tid + k * NT;
        out[tbase + idx] = lds[pad(idx)];
    }
}

// Row pass fused with the key inner product (f64 primes): for one limb and tile, every digit
// not owning the limb is row-transformed from its column-pass output, the owning digit's limb
// is read as it is (NTT form), and acc_b / acc_a = sum_j digit_j * key part j (b / a) are
// written; the transformed digits never reach memory (keyswitch.rs:254-303).
template <u32 M>
CKH_DEV void rows_ip(const FpAr& ar, const RowIpJob& J, u64* lds, u32 log_n, u32 tile) {
    const u64 tbase = (u64)tile << TL;
    const double q = ar.q, qinv = ar.qinv;
    double sb[16], sa[16];
#pragma unroll
    for (u32 k = 0; k < 16; ++k) sb[k] = sa[k] = 0.0;
    for (u32 j = 0; j < J.ndig; ++j) {
        // The thread index, opaque to the compiler per digit: otherwise it hoists every address
        // derived from it (16 words x the row pass's and the products' operands) out of the
        // digit loop and keeps them live throughout, 244 VGPRs (2 waves per SIMD) instead of 166
        // (3 waves); recomputing them costs a few integer ops per digit.
        u32 tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const auto kb = glb(J.key + (u64)(2 * j) * J.key_poly + tbase);
        const auto ka = glb(J.key + (u64)(2 * j + 1) * J.key_poly + tbase);
        if (j == J.own_j) {
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                const double x = u2d(glb(J.own)[tbase + tid + k * NT]);
                const double b = u2d(kb[tid + k * NT]), a = u2d(ka[tid + k * NT]);
                sb[k] += fmulmod(x, b, b * qinv, q);
                sa[k] += fmulmod(x, a, a * qinv, q);
            }
        } else {
            fwd_rows_to_lds<M, true>(ar, J.mid + j * J.mid_stride, lds, log_n, tile, tid);
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                const double x = __longlong_as_double((long long)lds[pad(tid + k * NT)]);
                const double b = u2d(kb[tid + k * NT]), a = u2d(ka[tid + k * NT]);
                sb[k] += fmulmod(x, b, b * qinv, q);
                sa[k] += fmulmod(x, a, a * qinv, q);
            }
            __syncthreads();  // the next digit's pass reuses the LDS
        }
        if ((j & 3) == 3) {
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                sb[k] = freduce(sb[k], q, qinv);
                sa[k] = freduce(sa[k], q, qinv);
            }
        }
    }
    const u32 tid = threadIdx.x;
#pragma unroll
    for (u32 k = 0; k < 16; ++k) {
        glb(J.acc_b)[tbase + tid + k * NT] = d2u(fcanon(sb[k], q, qinv));
        glb(J.acc_a)[tbase + tid + k * NT] = d2u(fcanon(sa[k], q, qinv));
    }
}

template <u32 M, class Ar>
CKH_DEV void inv_rows(const Ar& ar, const u64* __restrict__ in_p, u64* __restrict__ out_p, u64* lds, u32 log_n,
                      u32 tile, u32 tid = threadIdx.x) {
    const auto in = glb(in_p);
    const auto out = glb(out_p);
    constexpr u32 R2 = M - 4, U = 4 - R2, TRL = TL - M, JM = (1u << R2) - 1;
    const u32 o = tid & JM, rr = tid >> R2;
    const u32 row = (tile << TRL) + rr;
    const u64 tbase = (u64)tile << TL;
#pragma unroll
    for (u32 k = 0; k < 16; ++k) {
        const u32 idx = tid + k * NT;
        lds[pad(idx)] = in[tbase + idx];
    }
    __syncthreads();
    typename Ar::V v[16];
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 c = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        v[e] = ar.from_canon(lds[pad((rr << M) + c)]);
    }
    // Round 1: c = 16 o + e; stage u uses tab[2^(logN-1-u) + row 2^(M-1-u) + o 2^(3-u) + m],
    // m = e >> (u + 1).
    rows_round1_inv<M, Ar>(ar, v, log_n - M, row, o);
    __syncthreads();
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 c = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        lds[pad((rr << M) + c)] = ar.to_mid(v[e]);
    }
    __syncthreads();
#pragma unroll
    for (u32 k = 0; k < 16; ++k) v[k] = ar.from_mid(lds[pad((rr << M) + o + (k << R2))]);
    // Round 2: c = o + (k << R2) with o < 2^R2, so c >> (u + 1) = k >> (u + 1 - R2): stage u
    // uses tab[2^(logN-1-u) + row 2^(M-1-u) + m], m < 2^(M-1-u).
    rows_round2_inv<M, Ar>(ar, v, log_n - M, row);
    const u64 base = (u64)row << M;
#pragma unroll
    for (u32 k = 0; k < 16; ++k) out[base + o + (k << R2)] = ar.to_mid(v[k]);
}

template <u32 M, class Ar, bool COH = false>
CKH_DEV void inv_cols(const Ar& ar, u64* __restrict__ io_p, u64* lds, u32 log_n, u32 tile, u32 tid = threadIdx.x) {
    const auto io = glb(io_p);
    constexpr u32 R2 = M - 4, U = 4 - R2, TCL = TL - M, TC = 1u << TCL, JM = (1u << R2) - 1;
    const u32 B = log_n - M;
    const u32 cc = tid & (TC - 1), o = tid >> TCL;
    const u32 c0 = tile * TC;
    typename Ar::V v[16];
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        const u64 idx = ((u64)r << B) + c0 + cc;
        v[e] = ar.from_mid(COH ? load_coherent(io_p + idx) : io[idx]);
    }
#pragma unroll
    for (u32 u = 0; u < R2; ++u) {
#pragma unroll
        for (u32 e = 0; e < 16; ++e) {
            if (e & (1u << u)) continue;
            const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
            ar.gs(v[e], v[e + (1u << u)], ar.tw((1u << (log_n - 1 - B - u)) + (r >> (u + 1))));
        }
    }
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        lds[pad((r << TCL) + cc)] = ar.to_mid(v[e]);
    }
    __syncthreads();
#pragma unroll
    for (u32 k = 0; k < 16; ++k) v[k] = ar.from_mid(lds[pad(((o + (k << R2)) << TCL) + cc)]);
#pragma unroll
    for (u32 u = R2; u < M; ++u) {
        const u32 d = 1u << (u - R2);
#pragma unroll
        for (u32 k = 0; k < 16; ++k) {
            if (k & d) continue;
            if (u == M - 1) {
                ar.gs_final(v[k], v[k + d]);
            } else {
                const u32 r = o + (k << R2);
                ar.gs(v[k], v[k + d], ar.tw((1u << (log_n - 1 - B - u)) + (r >> (u + 1))));
            }
        }
    }
#pragma unroll
    for (u32 k = 0; k < 16; ++k) io[((u64)(o + (k << R2)) << B) + c0 + cc] = ar.canon_out(v[k]);
}

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

template <u32 M>
__global__ __launch_bounds__(NT) void k_inv_cols(const LimbRef* __restrict__ refs, const PrimeConst* __restrict__ pcs,
                                                 const Tw* __restrict__ tw, const FTw* __restrict__ twf,
                                                 const FTw* __restrict__ rt, u32 log_n,
                                                 u32 nrefs) {
    __shared__ u64 lds[LDS_WORDS];
    CKH_AR_DISPATCH((inv_cols<M>(ar, ref.dst, lds, log_n, tile)))
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

template <u32 M>
__global__ __launch_bounds__(NT) void k_rows_ip(const RowIpJob* __restrict__ jobs, const PrimeConst* __restrict__ pcs,
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
                       const FinishJob* fins = nullptr) {
    dim3 grid((1u << (t.log_n - fast::TL)) * nrefs);
    if (fwd && cols)
        hipLaunchKernelGGL(fast::k_fwd_cols<M>, grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf, t.log_n, nrefs);
    else if (fwd && fins)
        hipLaunchKernelGGL((fast::k_fwd_rows<M, true>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf,
                           t.log_n, nrefs, fins);
    else if (fwd)
        hipLaunchKernelGGL((fast::k_fwd_rows<M, false>), grid, dim3(fast::NT), 0, s, refs, t.pc, t.fwd, t.fwd_f, t.rf,
                           t.log_n, nrefs, fins);
    else if (cols)
        hipLaunchKernelGGL(fast::k_inv_cols<M>, grid, dim3(fast::NT), 0, s, refs, t.pc, t.inv, t.inv_f, t.ri, t.log_n, nrefs);
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
                     const FinishJob* fins = nullptr) {
    switch (m) {
        case 6: launch_one<6>(fwd, cols, t, refs, nrefs, s, fins); break;
        case 7: launch_one<7>(fwd, cols, t, refs, nrefs, s, fins); break;
        default: launch_one<8>(fwd, cols, t, refs, nrefs, s, fins); break;
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

void rows_ip_launch(const NttTables& t, const RowIpJob* jobs, u32 njobs, hipStream_t s) {
    if (njobs == 0) return;
    dim3 grid((1u << (t.log_n - fast::TL)) * njobs);
    switch (t.log_n / 2) {
        case 6: hipLaunchKernelGGL(fast::k_rows_ip<6>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        case 7: hipLaunchKernelGGL(fast::k_rows_ip<7>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
        default: hipLaunchKernelGGL(fast::k_rows_ip<8>, grid, dim3(fast::NT), 0, s, jobs, t.pc, t.fwd_f, t.rf, t.log_n, njobs); break;
    }
}

void ntt_fast_inverse(const NttTables& t, const LimbRef* refs, u32 nrefs, hipStream_t s) {
    const u32 B = t.log_n / 2, A = t.log_n - B;
    launch_m(B, false, false, t, refs, nrefs, s);
    launch_m(A, false, true, t, refs, nrefs, s);
}

}  // namespace ckh
