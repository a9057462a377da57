// The key inner product per coefficient (the body of k_inner_product) and the NTT-form Galois
// index, written so that a host program can include them too (the CPU replay in tests/emu). The
// includer has included arith.h and defines glb(p) and brev32(x) (32-bit bit reversal).
#pragma once

#include "arith.h"

namespace ckh {

// One limb of a key-switch inner product: digit j's limb at dig + j * dig_stride (or own for
// digit own_j when own is set), read through the Galois permutation when galois != 0; key part
// j's b limb at key + 2 j key_poly, its a limb key_poly further; acc_b / acc_a receive the sums.
struct IpJob {
    const u64* dig;
    u64 dig_stride;
    const u64* key;
    u64 key_poly;
    u64* acc_b;
    u64* acc_a;
    u64 galois;
    const u64* own;
    u32 ndig, prime, own_j, pad;
};

// NTT-form Galois permutation: sigma_g(a)[i] = a[perm(i)],
// perm(i) = bitrev(((g * (2 bitrev(i) + 1)) mod 2N - 1) / 2)
// (crates/ckks-train-core/src/rns.rs:5-17,147-157), computed inline instead of read from a table.
CKH_DEV u32 galois_index(u32 i, u64 g, u32 log_n) {
    const u32 rev = (brev32(i) >> (32 - log_n)) * 2 + 1;
    const u32 raw = (u32)(((g * rev) >> 1) & ((1u << log_n) - 1));
    return brev32(raw) >> (32 - log_n);
}

// acc_b = sum_j d_j * b_j and acc_a = sum_j d_j * a_j per limb of Q_l * P
// (crates/ckks-train-ckks/src/keyswitch.rs:254-303; residues are unique, so any exact summation
// order gives the reference's words). Primes below 2^50 sum f64 modular products (reduced every
// four digits); larger primes sum exact 128-bit products and reduce once. Digit limbs may be read
// through the Galois permutation (hoisted rotations, evaluator.rs:596-604).
// Digits are taken IP_GROUP at a time, every word of a group loaded before any product, so the
// loads of a group are in flight together.
constexpr u32 IP_GROUP = 4;

CKH_DEV void ip_coeff(const IpJob& J, const PrimeConst& C, u32 log_n, u32 c) {
    const u32 cs = J.galois ? galois_index(c, J.galois, log_n) : c;
    const auto key = glb(J.key);
    u64 dv[IP_GROUP], bv[IP_GROUP], av[IP_GROUP];
    auto load_group = [&](u32 j0) {
#pragma unroll
        for (u32 u = 0; u < IP_GROUP; ++u) {
            const u32 j = j0 + u;
            if (j < J.ndig) {
                dv[u] = glb((J.own && j == J.own_j) ? J.own : J.dig + j * J.dig_stride)[cs];
                bv[u] = key[(2 * j) * J.key_poly + c];
                av[u] = key[(2 * j + 1) * J.key_poly + c];
            }
        }
    };
    if (C.q < FP_PRIME_BOUND) {
        const double q = C.qf, qinv = C.qinv;
        double sb = 0.0, sa = 0.0;
        for (u32 j0 = 0; j0 < J.ndig; j0 += IP_GROUP) {
            load_group(j0);
#pragma unroll
            for (u32 u = 0; u < IP_GROUP; ++u) {
                if (j0 + u < J.ndig) {
                    const double d = u2d(dv[u]), b = u2d(bv[u]), a = u2d(av[u]);
                    sb += fmulmod(d, b, b * qinv, q);
                    sa += fmulmod(d, a, a * qinv, q);
                }
            }
            // four products (|.| < 0.7 q each) since the last reduction
            sb = freduce(sb, q, qinv);
            sa = freduce(sa, q, qinv);
        }
        glb(J.acc_b)[c] = d2u(fcanon(sb, q, qinv));
        glb(J.acc_a)[c] = d2u(fcanon(sa, q, qinv));
        return;
    }
    Dot60 sb, sa;
    for (u32 j0 = 0; j0 < J.ndig; j0 += IP_GROUP) {
        load_group(j0);
#pragma unroll
        for (u32 u = 0; u < IP_GROUP; ++u) {
            if (j0 + u < J.ndig) {
                sb.add(dv[u], bv[u]);
                sa.add(dv[u], av[u]);
            }
        }
    }
    u64 hi, lo;
    sb.get(hi, lo);
    glb(J.acc_b)[c] = reduce128(hi, lo, C);
    sa.get(hi, lo);
    glb(J.acc_a)[c] = reduce128(hi, lo, C);
}

}  // namespace ckh
