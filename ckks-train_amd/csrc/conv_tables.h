// Host construction of the conversion tables the conversion kernels read (BaseConverter::new,
// crates/ckks-train-core/src/baseconv.rs:95-160: (Q/q_j)^-1 mod q_j per source, (Q/q_j) mod r,
// Q mod r and (Q-1)/2 mod r per target), in the forms of every arithmetic path. Used by the
// context (which copies them to the device) and by the CPU replay of the conversion. The
// includer has included conv_arith.h first.
#pragma once

#include <utility>
#include <vector>

#include "hostmath.h"

namespace ckh {

struct HostConvTables {
    std::vector<ConvSrc> srcs;
    std::vector<ConvTgt> tgts;
    std::vector<ConvPlan> plans;

    // A plan from the primes `src` (indices into primes; input limbs limb0 + j) to the targets
    // `tg` (prime index, output limb). With share_tgt_first the plan's targets are the first
    // tg.size() of the list already at shared_first. Returns the plan index.
    u32 add_plan(const std::vector<u64>& primes, const std::vector<u32>& src,
                 const std::vector<std::pair<u32, u32>>& tg, u32 limb0, bool share_tgt_first, u32 shared_first) {
        ConvPlan pl{};
        pl.src_first = (u32)srcs.size();
        pl.nsrc = (u32)src.size();
        pl.src_limb0 = limb0;
        pl.exact = src.size() > 1;
        std::vector<u64> sq;
        for (u32 s : src) sq.push_back(primes[s]);
        for (size_t j = 0; j < src.size(); ++j) {
            const u64 q = sq[j];
            u64 qh = 1 % q;
            for (size_t i = 0; i < src.size(); ++i)
                if (i != j) qh = hm::mul(qh, sq[i] % q, q);
            ConvSrc cs{};
            cs.q = q;
            cs.h = q >> 1;
            cs.qhinv = hm::inv(qh, q);
            cs.qhinv_s = hm::shoup(cs.qhinv, q);
            cs.inv = 1.0 / (double)q;
            cs.fp = q < FP_PRIME_BOUND;
            cs.qf = (double)q;
            cs.qhinv_f = (double)cs.qhinv;
            cs.qhinv_q = (double)cs.qhinv / (double)q;
            srcs.push_back(cs);
        }
        if (share_tgt_first) {
            pl.tgt_first = shared_first;
        } else {
            pl.tgt_first = (u32)tgts.size();
            for (auto [prime, limb] : tg) {
                const u64 r = primes[prime];
                ConvTgt t{};
                t.prime = prime;
                t.limb = limb;
                t.qmod = hm::prod_mod(sq.data(), sq.size(), r);
                t.qmod_s = hm::shoup(t.qmod, r);
                t.hmod = hm::half_mod(sq.data(), sq.size(), r);
                bool srcs_fp = true;
                for (u64 sqj : sq) srcs_fp = srcs_fp && sqj < FP_PRIME_BOUND;
                t.mode = r >= FP_PRIME_BOUND ? TGT_INT : (srcs_fp ? TGT_FP : TGT_SPLIT);
                t.rf = (double)r;
                t.rinv = 1.0 / (double)r;
                t.qmod_f = (double)t.qmod;
                t.qmod_q = (double)t.qmod / (double)r;
                t.hmod_f = (double)t.hmod;
                if (r < FP_PRIME_BOUND) {
                    for (auto [sh, dst] : {std::pair<u32, FTw*>{20, &t.p20}, {40, &t.p40}, {65, &t.p65}}) {
                        const u64 v = hm::pow(2, sh, r);
                        *dst = {(double)v, (double)v / (double)r};
                    }
                }
                for (size_t j = 0; j < src.size(); ++j) {
                    u64 qh = 1 % r;
                    for (size_t i = 0; i < src.size(); ++i)
                        if (i != j) qh = hm::mul(qh, sq[i] % r, r);
                    t.qhat[j] = qh;
                    t.qhat_f[j] = (double)qh;
                    t.qhat_q[j] = (double)qh / (double)r;
                    const u64 qh30 = hm::mul(qh, (1ull << 30) % r, r);
                    t.qhat30_f[j] = (double)qh30;
                    t.qhat30_q[j] = (double)qh30 / (double)r;
                    for (u32 b = 0; b < 2; ++b) t.half[b][j] = (double)((qh >> (25 * b)) & ((1u << 25) - 1));
                }
                tgts.push_back(t);
            }
        }
        pl.ntgt = (u32)tg.size();
        plans.push_back(pl);
        return (u32)plans.size() - 1;
    }
};

}  // namespace ckh
