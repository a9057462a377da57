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
    // the matrix-core form (build_mf): per plan, the A tiles (1024 bytes each), per target
    std::vector<MfPlan> mf_plans;
    std::vector<int8_t> mf_a;
    std::vector<MfTgt> mf_tgt;

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
            cs.prime = src[j];
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
                t.nq = (r - t.qmod % r) % r;
                t.nh = (r - t.hmod % r) % r;
                t.nq_c = t.nq > r / 2 ? -(double)(r - t.nq) : (double)t.nq;
                for (u32 b = 0; b < 2; ++b) {
                    t.nq_half[b] = (double)((t.nq >> (25 * b)) & ((1u << 25) - 1));
                    t.nh_half[b] = (double)((t.nh >> (25 * b)) & ((1u << 25) - 1));
                }
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

    // The YForm of every source (same index as srcs), from each prime's inverse-NTT constants.
    std::vector<YForm> yform(const std::vector<PrimeConst>& pc) const {
        std::vector<YForm> out(srcs.size());
        for (size_t i = 0; i < srcs.size(); ++i) {
            const ConvSrc& cs = srcs[i];
            const u64 q = cs.q;
            YForm& y = out[i];
            y.n = hm::mul(pc[cs.prime].n_inv, cs.qhinv, q);
            y.wn = hm::mul(pc[cs.prime].w_n_inv, cs.qhinv, q);
            y.h = hm::mul(cs.h, cs.qhinv, q);
            y.n_s = hm::shoup(y.n, q);
            y.wn_s = hm::shoup(y.wn, q);
            y.n_f = {(double)y.n, (double)y.n / (double)q};
            y.wn_f = {(double)y.wn, (double)y.wn / (double)q};
            y.h_f = (double)y.h;
        }
        return out;
    }

    // Balanced base-256 digits of v < 2^63: v = sum_p d_p 256^p, d_p in [-128, 127], 8 digits.
    static void balanced_digits(u64 v, int8_t d[8]) {
        for (u32 p = 0; p < 8; ++p) {
            int b = (int)(v & 255);
            if (b >= 128) b -= 256;
            d[p] = (int8_t)b;
            v = (v - (u64)(int64_t)b) >> 8;  // exact: v - b is a multiple of 256
        }
    }

    // The matrix-core form of every conversion plan (MfPlan): A tiles in the lane order of
    // v_mfma_i32_32x32x32_i8 (lane l holds row l % 32, k = 16 (l / 32) + i for byte i; row
    // m = 16 (u % 2) + 8 (p / 4) + 4 (u / 2) + p % 4 for target u of the group and digit p, so
    // that after the product lane l holds all eight digits of targets 2 (l / 32) and 2 (l / 32) + 1
    // for coefficient l % 32; k = 8 j + b for source j, byte b), and the per-target recombination
    // constants.
    // Returns whether every plan has one (the matrix-core kernel may run).
    bool build_mf(const std::vector<u64>& primes, u64 n) {
        std::vector<MfPlan>& mp = mf_plans;
        std::vector<int8_t>& a = mf_a;
        std::vector<MfTgt>& mt = mf_tgt;
        mp.assign(plans.size(), MfPlan{});
        a.clear();
        mt.clear();
        bool all = n >= 256;
        for (size_t pi = 0; pi < plans.size(); ++pi) {
            const ConvPlan& P = plans[pi];
            MfPlan& F = mp[pi];
            F = {0, 0, 0, 0};
            if (P.nsrc == 0 || P.nsrc > 8 || n < 256) {
                all = false;
                continue;
            }
            // targets by recombination path, each list padded to a multiple of 4
            std::vector<int> order;
            for (int fp = 1; fp >= 0; --fp) {
                u32 cnt = 0;
                for (u32 t = 0; t < P.ntgt; ++t)
                    if ((primes[tgts[P.tgt_first + t].prime] < FP_PRIME_BOUND) == (fp == 1)) order.push_back((int)t), ++cnt;
                while (cnt % 4) order.push_back(-1), ++cnt;
            }
            const u32 ng = (u32)order.size() / 4;
            if (ng > MF_MAX_GROUPS) {
                all = false;
                continue;
            }
            F.ngroups = ng;
            F.kb = (P.nsrc + 3) / 4;
            F.a_off = (u32)(a.size() / 1024);
            F.t_off = (u32)mt.size();
            a.resize(a.size() + (size_t)ng * F.kb * 1024, 0);
            int8_t* A = a.data() + (size_t)F.a_off * 1024;
            for (u32 g = 0; g < ng; ++g)
                for (u32 u = 0; u < 4; ++u) {
                    const int ti = order[4 * g + u];
                    MfTgt T{};
                    T.limb = ~0u;
                    if (ti >= 0) {
                        const ConvTgt& ct = tgts[P.tgt_first + ti];
                        const u64 r = primes[ct.prime];
                        // W'_(j,b) = 2^(8b) c_j mod r and its digits into the A tiles; the byte offset
                        // (+128 per byte) contributes 128 sum W'.
                        u64 bsum = 0;
                        for (u32 j = 0; j < P.nsrc; ++j)
                            for (u32 b = 0; b < 8; ++b) {
                                const u64 w = hm::mul(ct.qhat[j], hm::pow(256 % r, b, r), r);
                                bsum = (bsum + w) % r;
                                int8_t d[8];
                                balanced_digits(w, d);
                                const u32 k = 8 * j + b, kb = k / 32, kk = k % 32;
                                for (u32 p = 0; p < 8; ++p) {
                                    const u32 m = 16 * (u % 2) + 8 * (p / 4) + 4 * (u / 2) + p % 4;
                                    const u32 lane = m + 32 * (kk / 16), byte = kk % 16;
                                    A[((size_t)(g * F.kb + kb) * 64 + lane) * 16 + byte] = d[p];
                                }
                            }
                        const u64 bias = hm::mul(128 % r, bsum, r);
                        T.limb = ct.limb;
                        T.fp = r < FP_PRIME_BOUND;
                        T.rf = (double)r;
                        T.rinv = 1.0 / (double)r;
                        const u64 c32 = hm::pow(2, 32, r);
                        T.c32_f = (double)c32;
                        T.c32_q = (double)c32 / (double)r;
                        T.bias_f = (double)((bias + r - ct.hmod % r) % r);
                        T.nq_c = ct.nq_c;
                        T.q = r;
                        T.nq = ct.nq;
                        T.corr = (u64)(((hm::u128)bias + r - ct.hmod % r) % r);
                        T.nq_f = (double)T.nq;
                        T.corr_f = (double)T.corr;
                    }
                    mt.push_back(T);
                }
        }
        return all;
    }
};

}  // namespace ckh
