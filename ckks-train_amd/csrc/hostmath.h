// Host-side modular arithmetic for building device tables: prime search, minimal primitive
// roots, inverses and Shoup companions. Follows crates/ckks-train-core/src/{modulus,primes,ntt}.rs
// (the same primes and the same psi, so the device tables equal the CPU path's).
#pragma once
#include <stdint.h>

#include <vector>

namespace ckh {
namespace hm {

typedef unsigned __int128 u128;

inline uint64_t mul(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

inline uint64_t pow(uint64_t a, uint64_t e, uint64_t q) {
    uint64_t acc = 1 % q;
    a %= q;
    while (e) {
        if (e & 1) acc = mul(acc, a, q);
        a = mul(a, a, q);
        e >>= 1;
    }
    return acc;
}

// Inverse by the extended Euclidean algorithm; 0 when gcd != 1 (modulus.rs:127-139).
inline uint64_t inv(uint64_t a, uint64_t q) {
    __int128 r0 = q, r1 = a % q, t0 = 0, t1 = 1;
    while (r1) {
        __int128 k = r0 / r1, t = r0 - k * r1;
        r0 = r1;
        r1 = t;
        t = t0 - k * t1;
        t0 = t1;
        t1 = t;
    }
    if (r0 != 1) return 0;
    __int128 v = t0 % (__int128)q;
    if (v < 0) v += q;
    return (uint64_t)v;
}

inline uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)(w % q) << 64) / q); }

// Deterministic Miller-Rabin, bases 2..37 (primes.rs:24-53).
inline bool is_prime(uint64_t n) {
    static const uint64_t B[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : B) {
        if (n == b) return true;
        if (n % b == 0) return false;
    }
    unsigned s = __builtin_ctzll(n - 1);
    uint64_t d = (n - 1) >> s;
    for (uint64_t b : B) {
        uint64_t x = pow(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (unsigned r = 1; r < s; ++r) {
            x = mul(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp) return false;
    }
    return true;
}

// The `count` largest primes 2^(bits-1) < q < 2^bits with q = 1 mod 2n, descending
// (primes.rs:67-112). Returns false when there are not enough.
inline bool ntt_primes(unsigned bits, uint64_t n, uint64_t count, std::vector<uint64_t>& out) {
    out.clear();
    if (bits < 2 || bits > 60 || n == 0 || (n & (n - 1))) return false;
    u128 step = (u128)n * 2, lo = (u128)1 << (bits - 1), hi = (u128)1 << bits;
    if (step >= hi) return count == 0;
    u128 q = (hi - 2) / step * step + 1;
    while (out.size() < count) {
        if (q <= lo) return false;
        if (is_prime((uint64_t)q)) out.push_back((uint64_t)q);
        if (q < step) return false;
        q -= step;
    }
    return true;
}

// Minimal primitive 2n-th root of unity mod prime q (ntt.rs:66-82).
inline uint64_t min_root(uint64_t n, uint64_t q) {
    uint64_t e = (q - 1) / (2 * n), c = 0;
    for (uint64_t g = 2;; ++g) {
        c = pow(g, e, q);
        if (pow(c, n, q) == q - 1) break;
    }
    uint64_t c2 = mul(c, c, q), cur = c, best = c;
    for (uint64_t i = 1; i < n; ++i) {
        cur = mul(cur, c2, q);
        if (cur < best) best = cur;
    }
    return best;
}

inline uint32_t bitrev(uint32_t i, unsigned bits) {
    uint32_t r = 0;
    for (unsigned b = 0; b < bits; ++b) r |= ((i >> b) & 1u) << (bits - 1 - b);
    return r;
}

// (prod of factors) mod m.
inline uint64_t prod_mod(const uint64_t* f, size_t k, uint64_t m) {
    uint64_t acc = 1 % m;
    for (size_t i = 0; i < k; ++i) acc = mul(acc, f[i] % m, m);
    return acc;
}

// (D - 1)/2 mod m for D = prod of factors (odd), m odd (baseconv.rs:238-243).
inline uint64_t half_mod(const uint64_t* f, size_t k, uint64_t m) {
    uint64_t d = prod_mod(f, k, m);
    uint64_t dm1 = d >= 1 % m ? d - 1 % m : d + m - 1 % m;
    return mul(dm1, (m + 1) / 2, m);
}

}  // namespace hm
}  // namespace ckh
