// Host construction of the per-prime NTT tables (NttTable::new, crates/ckks-train-core/src/
// ntt.rs:66-122: psi the minimal primitive 2N-th root, tables in bit-reversed order) in the
// forms the kernels read: Shoup pairs for the integer path, (w, w/q) f64 pairs for primes below
// 2^50, the per-prime constants, and the row factors of the register-radix row pass. Used by the
// context (which copies them to the device) and by the CPU replay of the kernels. The includer
// has included the arithmetic types (common.h, or the replay's prelude) first.
#pragma once

#include <vector>

#include "hostmath.h"

namespace ckh {

struct HostNttTables {
    std::vector<PrimeConst> pc;
    std::vector<Tw> fwd, inv;        // [prime][N]
    std::vector<FTw> fwd_f, inv_f;   // [prime][N]
    std::vector<FTw> rf, ri;         // [prime][2^A rows][ROW_FACTOR_SLOTS] (register-radix N only)
};

inline HostNttTables build_ntt_tables(const std::vector<u64>& primes, u32 log_n, bool row_factors) {
    const u64 n = 1ull << log_n;
    const u32 np_all = (u32)primes.size();
    HostNttTables t;
    t.pc.resize(np_all);
    t.fwd.resize((size_t)np_all * n);
    t.inv.resize((size_t)np_all * n);
    t.fwd_f.resize((size_t)np_all * n);
    t.inv_f.resize((size_t)np_all * n);
    for (u32 k = 0; k < np_all; ++k) {
        const u64 q = primes[k];
        const u64 psi = hm::min_root(n, q), psi_inv = hm::inv(psi, q), n_inv = hm::inv(n % q, q);
        u64 p1 = 1, p2 = 1;
        for (u64 i = 0; i < n; ++i) {
            const u32 j = hm::bitrev((u32)i, log_n);
            t.fwd[(size_t)k * n + j] = {p1, hm::shoup(p1, q)};
            t.inv[(size_t)k * n + j] = {p2, hm::shoup(p2, q)};
            t.fwd_f[(size_t)k * n + j] = {(double)p1, (double)p1 / (double)q};
            t.inv_f[(size_t)k * n + j] = {(double)p2, (double)p2 / (double)q};
            p1 = hm::mul(p1, psi, q);
            p2 = hm::mul(p2, psi_inv, q);
        }
        const u64 iroot1 = n >= 2 ? t.inv[(size_t)k * n + 1].w : 1;
        PrimeConst& c = t.pc[k];
        c.q = q;
        c.n_inv = n_inv;
        c.n_inv_s = hm::shoup(n_inv, q);
        c.w_n_inv = hm::mul(iroot1, n_inv, q);
        c.w_n_inv_s = hm::shoup(c.w_n_inv, q);
        c.one_s = hm::shoup(1, q);
        c.two64 = (u64)(((hm::u128)1 << 64) % q);
        c.two64_s = hm::shoup(c.two64, q);
        c.qf = (double)q;
        c.qinv = 1.0 / (double)q;
        c.n_inv_f = {(double)n_inv, (double)n_inv / (double)q};
        c.w_n_inv_f = {(double)c.w_n_inv, (double)c.w_n_inv / (double)q};
    }
    if (row_factors) {
        const u32 M = log_n / 2, A = log_n - M, rows = 1u << A;
        t.rf.resize((size_t)np_all * rows * ROW_FACTOR_SLOTS);
        t.ri.resize(t.rf.size());
        for (u32 k = 0; k < np_all; ++k)
            for (u32 row = 0; row < rows; ++row)
                for (u32 s = 0; s < M; ++s) {
                    const size_t at = ((size_t)k * rows + row) * ROW_FACTOR_SLOTS + s;
                    const size_t src = (size_t)k * n + (1u << (A + s)) + ((size_t)row << s);
                    t.rf[at] = t.fwd_f[src];
                    t.ri[at] = t.inv_f[src];
                }
    }
    return t;
}

}  // namespace ckh
