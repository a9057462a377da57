// Inner product, pointwise, automorphism and seeded-fill launchers.
#pragma once
#include "common.h"
#include "ip_arith.h"

namespace ckh {

// One limb of a key-switch inner product: acc_b = sum_j dig_j * key_b_j, acc_a = sum_j dig_j *
// key_a_j over ndig digits (digit j's limb at dig + j * dig_stride). The key limb of part j is
// at key + (2 j) * key_poly (b) and key + (2 j + 1) * key_poly (a). galois != 0 reads the digit
// through the NTT-form automorphism permutation (hoisted rotations). When own is set, digit
// own_j's limb (the digit's own, unconverted limb) is read from own instead.
// out (one limb, prime `prime`) <- sum over the job's terms [first, first + count) of x * w mod q:
// the BSGS block sums of plaintext products in one launch (MulSumTerm: two limb pointers).
struct MulSumJob {
    u64* out;
    u32 first, count, prime, pad;
};
struct MulSumTerm {
    const u64* x;
    const u64* w;
};
void mul_sum_launch(const MulSumJob* jobs, const MulSumTerm* terms, u32 njobs, const PrimeConst* pc, u32 log_n,
                    hipStream_t stream);

// out (one limb, prime `prime`) <- sum of the limbs xs[first, first + count) mod q.
void add_sum_launch(const MulSumJob* jobs, const u64* const* xs, u32 njobs, const PrimeConst* pc, u32 log_n,
                    hipStream_t stream);


void inner_product_launch(const IpJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n,
                          hipStream_t stream);

// ref.dst <- ref.src (one limb per descriptor).
// Grid-stride copy of bytes (a multiple of 16) with 16-byte loads and stores.
void stream_copy_launch(const void* in, void* out, u64 bytes, hipStream_t stream);
void copy_launch(const LimbRef* refs, u32 nrefs, u32 log_n, hipStream_t stream);

enum PwOp : u32 { PW_ADD = 0, PW_SUB = 1, PW_NEG = 2, PW_MUL = 3, PW_COPY = 4 };

// out limb i <- a op b over limbs [0, limbs) with prime index prime0 + i (Q primes).
void pointwise_launch(PwOp op, const u64* a, const u64* b, u64* out, u32 limbs, u32 prime0,
                      const PrimeConst* pc, u32 log_n, hipStream_t stream);

// out limb i <- (a limb i) * w_i mod q_i (Shoup pair (w_i, ws_i) at consts[2 i], consts[2 i + 1]),
// negated mod q_i when neg is set; Q primes 0..limbs.
void scalar_mul_launch(const u64* a, u64* out, u32 limbs, const u64* consts, bool neg, const PrimeConst* pc,
                       u32 log_n, hipStream_t stream);

// out limb i <- sigma_g(in limb i) in NTT form (index permutation, prime independent).
void automorphism_ntt_launch(const u64* in, u64* out, u32 limbs, u64 g, u32 log_n, hipStream_t stream);
// The same in coefficient form (a signed permutation: limb r of the layout reduced by its prime,
// Q primes for r <= level, then P primes).
void automorphism_coeff_launch(const u64* in, u64* out, u32 limbs, u32 level, u32 nq, u64 g, const PrimeConst* pc,
                               u32 log_n, hipStream_t stream);
// The same for many (limb, Galois element) jobs in one launch.
struct AutJob {
    const u64* in;
    u64* out;
    u64 galois;
};
void automorphism_batch_launch(const AutJob* jobs, u32 njobs, u32 log_n, hipStream_t stream);

// Seeded uniform residues: limb i of the layout gets prime primes_map(i) and generator stream
// stream0 + i; q_limbs limbs use Q primes 0.., the rest P primes (QpPoly layout).
void fill_uniform_launch(u64* out, u32 q_limbs, u32 p_limbs, u32 nq, u64 seed, u64 stream0,
                         const PrimeConst* pc, u32 log_n, hipStream_t stream);

// *out += sum over words of mix(word, index): an order-independent 64-bit digest of `words` words
// (for comparing large results without copying them to the host). *out must start at 0.
void digest_launch(const u64* data, u64 words, u64* out, hipStream_t stream);

}  // namespace ckh
