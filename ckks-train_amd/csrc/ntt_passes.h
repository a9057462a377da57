// The register-radix NTT passes for 2^12 <= N <= 2^16 (the tile schedule and its arithmetic),
// shared by the gfx950 kernels (ntt_fast.hip) and the CPU replay of those kernels (tests/emu).
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
//  * q >= 2^50 (IntAr): Harvey butterflies with 64-bit Shoup products, values in [0, 8q)
//    (forward: x reduced on alternate stages, every pass ending below 6q) or [0, 2q) (inverse).
//
// The includer provides, besides arith.h: glb(p) (a global-memory view of a pointer),
// load_coherent(p) (a load that sees other workgroups' completed stores), __syncthreads(),
// threadIdx.x and CKH_OPAQUE(x) (hides a value's origin from the optimizer). The LDS tile is a
// template parameter (a __shared__ array on the device).
#pragma once

#include "arith.h"

namespace ckh {
namespace fast {

constexpr u32 TL = 12;
constexpr u32 TILE = 1u << TL;
constexpr u32 NT = 256;
constexpr u32 LDS_WORDS = TILE + TILE / 16;

// Padded LDS index: one spare word per 16 spreads strided accesses over the banks.
CKH_DEV u32 pad(u32 a) { return a + (a >> 4); }

// ---------------------------------------------------------------- f64 arithmetic (q < 2^50)

struct FpAr {
    typedef double V;
    typedef FTw W;
    static constexpr bool F64 = true;
    const FTw* tab;
    double q, qinv;
    FTw ninv, wninv;
    const FTw* rt;    // this prime's row factors, [row][ROW_FACTOR_SLOTS]
    double add = 0.0; // added to the outputs of gs_final<true> (an integer below q)
    CKH_DEV V from_canon(u64 x) const { return u2d(x); }
    CKH_DEV V from_mid(u64 x) const { return bits_f64(x); }
    CKH_DEV u64 to_mid(V v) const { return f64_bits(v); }
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
    static constexpr bool REDUCE_LAST = false;
    CKH_DEV void ct_lazy(V& x, V& y, const W& w) const {
        const double t = fmulmod(y, w.w, w.wq, q);
        const double x0 = x;
        x = x0 + t;
        y = x0 - t;
    }
    CKH_DEV u64 fwd_final(V v) const { return d2u(fcanon(v, q, qinv)); }
    CKH_DEV u64 fwd_final_f(V v) const { return f64_bits(fcanon(v, q, qinv)); }
    CKH_DEV void gs(V& x, V& y, const W& w) const {
        const double s = x + y, d = x - y;
        x = freduce(s, q, qinv);
        y = fmulmod(d, w.w, w.wq, q);
    }
    // ADD: plus `add` (the products are below 0.7 q in magnitude, the sums below 1.7 q)
    template <bool ADD = false>
    CKH_DEV void gs_final(V& x, V& y) const {
        const double s = x + y, d = x - y;
        double a = fmulmod(s, ninv.w, ninv.wq, q), b = fmulmod(d, wninv.w, wninv.wq, q);
        if (ADD) {
            a = a + add;
            b = b + add;
        }
        x = fcanon(a, q, qinv);
        y = fcanon(b, q, qinv);
    }
    CKH_DEV u64 canon_out(V v) const { return d2u(v); }
};

// ---------------------------------------------------------------- integer arithmetic

struct IntAr {
    typedef u64 V;
    typedef Tw W;
    static constexpr bool F64 = false;
    const Tw* tab;
    u64 q;
    u64 n_inv, n_inv_s, w_n_inv, w_n_inv_s;
    u64 add = 0;  // added to the outputs of gs_final<true> (below q)
    CKH_DEV V from_canon(u64 x) const { return x; }
    CKH_DEV V from_mid(u64 x) const { return x; }
    CKH_DEV u64 to_mid(V v) const { return v; }
    CKH_DEV W tw(u32 idx) const { return tab[idx]; }
    template <u32 D>
    CKH_DEV void row_group(u32 full_base, u32, u32, u32, W* g) const {
#pragma unroll
        for (u32 m = 0; m < (1u << D); ++m) g[m] = tab[full_base + m];
    }
    // Forward butterflies over values below 8 q (q < 2^60, so below 2^63): the lazy product t is
    // below 2 q for any 64-bit y. Stages alternate as in the f64 path, but the last stage of each
    // pass always reduces (REDUCE_LAST): a pass takes inputs below 6 q, a lazy stage leaves them
    // below 8 q, a reduced one (x taken below 4 q) below 6 q, so every pass ends below 6 q.
    static constexpr bool REDUCE_LAST = true;
    CKH_DEV void ct(V& x, V& y, const W& w) const {
        const u64 four_q = 4 * q;
        const u64 xr = x >= four_q ? x - four_q : x;
        const u64 t = shoup_lazy(y, w.w, w.ws, q);
        x = xr + t;
        y = xr + 2 * q - t;
    }
    CKH_DEV void ct_lazy(V& x, V& y, const W& w) const {
        const u64 t = shoup_lazy(y, w.w, w.ws, q);
        const u64 x0 = x;
        x = x0 + t;
        y = x0 + 2 * q - t;
    }
    CKH_DEV u64 fwd_final_f(V v) const { return fwd_final(v); }  // not used (f64 primes only)
    CKH_DEV u64 fwd_final(V v) const {
        v = v >= 4 * q ? v - 4 * q : v;
        v = v >= 2 * q ? v - 2 * q : v;
        return v >= q ? v - q : v;
    }
    CKH_DEV void gs(V& x, V& y, const W& w) const {
        const u64 two_q = 2 * q;
        const u64 s = x + y, d = x + two_q - y;
        x = s >= two_q ? s - two_q : s;
        y = shoup_lazy(d, w.w, w.ws, q);
    }
    // ADD: plus `add` (the lazy products are below 2 q, so the sums below 3 q)
    template <bool ADD = false>
    CKH_DEV void gs_final(V& x, V& y) const {
        const u64 s = x + y, d = x + 2 * q - y;
        if (ADD) {
            x = csub(csub(shoup_lazy(s, n_inv, n_inv_s, q) + add, 2 * q), q);
            y = csub(csub(shoup_lazy(d, w_n_inv, w_n_inv_s, q) + add, 2 * q), q);
        } else {
            x = csub(shoup_lazy(s, n_inv, n_inv_s, q), q);
            y = csub(shoup_lazy(d, w_n_inv, w_n_inv_s, q), q);
        }
    }
    CKH_DEV u64 canon_out(V v) const { return v; }
};

// ---------------------------------------------------------------- row-pass twiddle groups
//
// In the row passes every thread needs, per stage, 2^D twiddles at consecutive table indices
// starting at a multiple of 2^D; Ar::group builds them (for the f64 path from one load).

// Forward butterfly of stage s of a pass of M stages (s is a constant after unrolling; M = 0:
// not the pass's last stage).
template <class Ar, u32 M = 0>
CKH_DEV void ct_stage(const Ar& ar, u32 s, typename Ar::V& x, typename Ar::V& y, const typename Ar::W& w) {
    if ((s & 1) || (Ar::REDUCE_LAST && s + 1 == M))
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
            ct_stage<Ar, M>(ar, s, v[e], v[e + (1u << p)], g[e >> (p + 1)]);
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

template <u32 M, class Ar, class Lds>
CKH_DEV void fwd_cols(const Ar& ar, const u64* __restrict__ in_p, u64* __restrict__ out_p, Lds lds, u32 log_n,
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
            ct_stage<Ar, M>(ar, s, v[e], v[e + (1u << p)], ar.tw((1u << s) + (r >> (M - s))));
        }
    }
#pragma unroll
    for (u32 e = 0; e < 16; ++e) {
        const u32 r = ((((o << U) | (e >> R2)) << R2) | (e & JM));
        out[((u64)r << B) + c0 + cc] = ar.to_mid(v[e]);
    }
}

// The row pass of one tile, its canonical outputs left in LDS in natural order (word i of the
// tile at pad(i)): as u64 residues, or (F64) as the f64 bits of the residue, or (RAW, f64 path)
// as the f64 bits of the last stage's signed outputs (|.| <= 2.6 q), not reduced.
// pre() runs after the last LDS stores, before the closing barrier (the caller's loads issued
// there are in flight while the workgroup meets).
struct NoPre {
    CKH_DEV void operator()() const {}
};
template <u32 M, bool F64, class Ar, bool COH = false, class Lds, bool RAW = false, class Pre = NoPre>
CKH_DEV void fwd_rows_to_lds(const Ar& ar, const u64* __restrict__ in_p, Lds lds, u32 log_n, u32 tile,
                             u32 tid = threadIdx.x, Pre pre = Pre{}) {
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
        lds[pad((rr << M) + c)] = RAW ? ar.to_mid(v[e]) : F64 ? ar.fwd_final_f(v[e]) : ar.fwd_final(v[e]);
    }
    pre();
    __syncthreads();
}

template <u32 M, class Ar, bool COH = false, class Lds>
CKH_DEV void fwd_rows(const Ar& ar, const u64* __restrict__ in, u64* __restrict__ out_p, Lds lds, u32 log_n,
                      const FinishJob* fin, u64 q, u32 tile, u32 tid = threadIdx.x) {
    const auto out = glb(out_p);
    const u64 tbase = (u64)tile << TL;
    if (fin) {
        // finish of rescale / mod_down: (a - t) * inv (+ add)
        const auto fa = glb(fin->a), fadd = glb(fin->add);
        const auto fout = glb(fin->out);
        const bool has_add = fin->add != nullptr;
        const u64 inv = fin->inv, inv_s = fin->inv_s;
        u64 av[16], dv[16];
        // a (and the addend) loaded once the row pass's registers are free, before its last barrier
        auto pre = [&] {
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                av[k] = fa[tbase + tid + k * NT];
                dv[k] = has_add ? (u64)fadd[tbase + tid + k * NT] : (u64)0;
            }
        };
        if constexpr (Ar::F64) {
            // in f64: the transform left unreduced (|.| <= 2.6 q), a - t below 3.6 q, one modular
            // product (|.| <= 1.2 q), the addend, one reduction
            fwd_rows_to_lds<M, false, Ar, COH, Lds, true>(ar, in, lds, log_n, tile, tid, pre);
            const FTw iv = fin->inv_f;
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                double v = fmulmod(u2d(av[k]) - bits_f64(lds[pad(tid + k * NT)]), iv.w, iv.wq, ar.q);
                if (has_add) v = v + u2d(dv[k]);
                fout[tbase + tid + k * NT] = d2u(fcanon(v, ar.q, ar.qinv));
            }
        } else {
            fwd_rows_to_lds<M, false, Ar, COH, Lds, false>(ar, in, lds, log_n, tile, tid, pre);
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                u64 v = shoup(sub_mod(av[k], lds[pad(tid + k * NT)], q), inv, inv_s, q);
                if (has_add) v = add_mod(v, dv[k], q);
                fout[tbase + tid + k * NT] = v;
            }
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

template <u32 M, class Ar, class Lds>
CKH_DEV void inv_rows(const Ar& ar, const u64* __restrict__ in_p, u64* __restrict__ out_p, Lds lds, u32 log_n,
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
    for (u32 k = 0; k < 16; ++k) st_nt(out, base + o + (k << R2), ar.to_mid(v[k]));
}

// ADD: the last stage adds ar.add to every output (the y form of a conversion's source limb).
template <u32 M, class Ar, bool COH = false, bool ADD = false, class Lds>
CKH_DEV void inv_cols(const Ar& ar, u64* __restrict__ io_p, Lds lds, u32 log_n, u32 tile, u32 tid = threadIdx.x) {
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
                ar.template gs_final<ADD>(v[k], v[k + d]);
            } else {
                const u32 r = o + (k << R2);
                ar.gs(v[k], v[k + d], ar.tw((1u << (log_n - 1 - B - u)) + (r >> (u + 1))));
            }
        }
    }
#pragma unroll
    for (u32 k = 0; k < 16; ++k) st_nt(io, ((u64)(o + (k << R2)) << B) + c0 + cc, ar.canon_out(v[k]));
}

// ---------------------------------------------------------------- row pass + inner product

// Row pass fused with the key inner product (f64 primes): for one limb and tile, every digit
// not owning the limb is row-transformed from its column-pass output, the owning digit's limb
// is read as it is (NTT form), and sb / sa = sum_j digit_j * key part j (b / a) are formed as
// signed f64 values (|.| <= q / 2 + 4 q); the transformed digits never reach memory
// (keyswitch.rs:254-303).
template <u32 M, class Lds>
CKH_DEV void rows_ip_sums(const FpAr& ar, const RowIpJob& J, Lds lds, u32 log_n, u32 tile, double (&sb)[16],
                          double (&sa)[16]) {
    const u64 tbase = (u64)tile << TL;
    const double q = ar.q, qinv = ar.qinv;
#pragma unroll
    for (u32 k = 0; k < 16; ++k) sb[k] = sa[k] = 0.0;
    for (u32 j = 0; j < J.ndig; ++j) {
        // The thread index, opaque to the compiler per digit: otherwise it hoists every address
        // derived from it (16 words x the row pass's and the products' operands) out of the
        // digit loop and keeps them live throughout, 244 VGPRs (2 waves per SIMD) instead of 166
        // (3 waves); recomputing them costs a few integer ops per digit.
        u32 tid = threadIdx.x;
        CKH_OPAQUE(tid);
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
            // the transform left unreduced (|x| <= 2.6 q: each product below q in magnitude); the
            // key words are loaded once the row pass's registers are free, before its last barrier
            u64 kbw[16], kaw[16];
            auto pre = [&] {
#pragma unroll
                for (u32 k = 0; k < 16; ++k) {
                    kbw[k] = kb[tid + k * NT];
                    kaw[k] = ka[tid + k * NT];
                }
            };
            fwd_rows_to_lds<M, true, FpAr, false, Lds, true>(ar, J.mid + j * J.mid_stride, lds, log_n, tile, tid, pre);
#pragma unroll
            for (u32 k = 0; k < 16; ++k) {
                const double x = bits_f64(lds[pad(tid + k * NT)]);
                const double b = u2d(kbw[k]), a = u2d(kaw[k]);
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
}

// The sums written as canonical residues to J.acc_b / J.acc_a.
template <u32 M, class Lds>
CKH_DEV void rows_ip(const FpAr& ar, const RowIpJob& J, Lds lds, u32 log_n, u32 tile) {
    const u64 tbase = (u64)tile << TL;
    double sb[16], sa[16];
    rows_ip_sums<M>(ar, J, lds, log_n, tile, sb, sa);
    const u32 tid = threadIdx.x;
#pragma unroll
    for (u32 k = 0; k < 16; ++k) {
        glb(J.acc_b)[tbase + tid + k * NT] = d2u(fcanon(sb[k], ar.q, ar.qinv));
        glb(J.acc_a)[tbase + tid + k * NT] = d2u(fcanon(sa[k], ar.q, ar.qinv));
    }
}

// mod_down's finish of one sum for one tile (baseconv.rs:459-511): the row pass of the converted
// limb z (column pass done) into LDS, then out = (sum - NTT(z)) * P^-1 (+ add). The LDS is free
// again on return.
// All in f64: the row pass leaves NTT(z) unreduced (|.| <= 2.6 q), the sum is reduced to q / 2,
// their difference (below 3.1 q) times P^-1 is one f64 modular product (|.| <= 1.1 q), the addend
// (below q) is added and the result reduced once.
template <u32 M, class Lds>
CKH_DEV void finish_sum(const FpAr& ar, const double (&s)[16], const u64* z, u64* out_p, const u64* add_p, FTw inv,
                        Lds lds, u32 log_n, u32 tile) {
    u32 tid = threadIdx.x;
    CKH_OPAQUE(tid);
    const u64 tbase = (u64)tile << TL;
    const double q = ar.q, qinv = ar.qinv;
    // the addend's words are loaded once the row pass's registers are free, before its last barrier
    u64 dv[16];
    auto pre = [&] {
        if (add_p) {
#pragma unroll
            for (u32 k = 0; k < 16; ++k) dv[k] = glb(add_p)[tbase + tid + k * NT];
        }
    };
    fwd_rows_to_lds<M, false, FpAr, false, Lds, true>(ar, z, lds, log_n, tile, tid, pre);
    const auto out = glb(out_p);
#pragma unroll
    for (u32 k = 0; k < 16; ++k) {
        const double d = freduce(s[k], q, qinv) - bits_f64(lds[pad(tid + k * NT)]);
        double v = fmulmod(d, inv.w, inv.wq, q);
        if (add_p) v = v + u2d(dv[k]);
        out[tbase + tid + k * NT] = d2u(fcanon(v, q, qinv));
    }
    __syncthreads();  // the next pass reuses the LDS
}

}  // namespace fast
}  // namespace ckh
