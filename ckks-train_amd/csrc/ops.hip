// Key-switch inner product, pointwise arithmetic, NTT-form automorphism and the seeded
// input generator, for gfx950.
#include <algorithm>

#include "ops.h"

namespace ckh {

// The key inner product of one coefficient (ip_arith.h: ip_coeff) per thread.
__global__ __launch_bounds__(256) void k_inner_product(const IpJob* __restrict__ jobs,
                                                       const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const IpJob J = jobs[blockIdx.y];
    ip_coeff(J, pc[J.prime], log_n, c);
}

void inner_product_launch(const IpJob* jobs, u32 njobs, const PrimeConst* pc, u32 log_n, hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_inner_product, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, pc, log_n);
}

__global__ __launch_bounds__(256) void k_copy(const LimbRef* __restrict__ refs, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const LimbRef r = refs[blockIdx.y];
    glb(r.dst)[c] = glb(r.src)[c];
}

void copy_launch(const LimbRef* refs, u32 nrefs, u32 log_n, hipStream_t stream) {
    if (nrefs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_copy, dim3((n + 255) / 256, nrefs), dim3(256), 0, stream, refs, log_n);
}

// Streaming copy of n16 16-byte words (a measurement of the HBM copy rate a kernel reaches).
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ in, uint4* __restrict__ out, u64 n16) {
    const u64 stride = (u64)gridDim.x * 256;
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride) out[i] = in[i];
}

void stream_copy_launch(const void* in, void* out, u64 bytes, hipStream_t stream) {
    const u64 n16 = bytes / 16;
    const u32 blocks = (u32)std::min<u64>((n16 + 255) / 256, 16384);
    if (blocks) hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, stream, (const uint4*)in, (uint4*)out, n16);
}

__global__ __launch_bounds__(256) void k_pointwise(u32 op, const u64* __restrict__ a, const u64* __restrict__ b,
                                                   u64* __restrict__ out, u32 prime0,
                                                   const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const u64 i = (u64)blockIdx.y * n + c;
    const PrimeConst C = pc[prime0 + blockIdx.y];
    u64 v;
    switch (op) {
        case PW_ADD: v = add_mod(a[i], b[i], C.q); break;
        case PW_SUB: v = sub_mod(a[i], b[i], C.q); break;
        case PW_NEG: v = sub_mod(0, a[i], C.q); break;
        case PW_MUL: v = mul_mod(a[i], b[i], C); break;
        default: v = a[i]; break;
    }
    out[i] = v;
}

void pointwise_launch(PwOp op, const u64* a, const u64* b, u64* out, u32 limbs, u32 prime0, const PrimeConst* pc,
                      u32 log_n, hipStream_t stream) {
    if (limbs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_pointwise, dim3((n + 255) / 256, limbs), dim3(256), 0, stream, (u32)op, a, b, out, prime0,
                       pc, log_n);
}

__global__ __launch_bounds__(256) void k_mul_sum(const MulSumJob* __restrict__ jobs,
                                                 const MulSumTerm* __restrict__ terms,
                                                 const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const MulSumJob J = jobs[blockIdx.y];
    const PrimeConst C = pc[J.prime];
    u64 acc = 0;
    for (u32 t = 0; t < J.count; ++t) {
        const MulSumTerm T = terms[J.first + t];
        acc = add_mod(acc, mul_mod(glb(T.x)[c], glb(T.w)[c], C), C.q);
    }
    glb(J.out)[c] = acc;
}

void mul_sum_launch(const MulSumJob* jobs, const MulSumTerm* terms, u32 njobs, const PrimeConst* pc, u32 log_n,
                    hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_mul_sum, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, terms, pc, log_n);
}

__global__ __launch_bounds__(256) void k_add_sum(const MulSumJob* __restrict__ jobs, const u64* const* __restrict__ xs,
                                                 const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const MulSumJob J = jobs[blockIdx.y];
    const u64 q = pc[J.prime].q;
    u64 acc = 0;
    for (u32 t = 0; t < J.count; ++t) acc = add_mod(acc, glb(xs[J.first + t])[c], q);
    glb(J.out)[c] = acc;
}

void add_sum_launch(const MulSumJob* jobs, const u64* const* xs, u32 njobs, const PrimeConst* pc, u32 log_n,
                    hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_add_sum, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, xs, pc, log_n);
}

__global__ __launch_bounds__(256) void k_scalar_mul(const u64* __restrict__ a, u64* __restrict__ out,
                                                    const u64* __restrict__ consts, u32 neg,
                                                    const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const u64 i = (u64)blockIdx.y * n + c;
    const u64 q = pc[blockIdx.y].q, w = consts[2 * blockIdx.y], ws = consts[2 * blockIdx.y + 1];
    u64 v = shoup(a[i], w, ws, q);
    if (neg) v = sub_mod(0, v, q);
    out[i] = v;
}

void scalar_mul_launch(const u64* a, u64* out, u32 limbs, const u64* consts, bool neg, const PrimeConst* pc,
                       u32 log_n, hipStream_t stream) {
    if (limbs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_scalar_mul, dim3((n + 255) / 256, limbs), dim3(256), 0, stream, a, out, consts, (u32)neg, pc,
                       log_n);
}

__global__ __launch_bounds__(256) void k_automorphism_ntt(const u64* __restrict__ in, u64* __restrict__ out, u64 g,
                                                          u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const u64 base = (u64)blockIdx.y * n;
    out[base + c] = in[base + galois_index(c, g, log_n)];
}

void automorphism_ntt_launch(const u64* in, u64* out, u32 limbs, u64 g, u32 log_n, hipStream_t stream) {
    if (limbs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_automorphism_ntt, dim3((n + 255) / 256, limbs), dim3(256), 0, stream, in, out, g, log_n);
}

// Coefficient form (crates/ckks-train-core/src/poly.rs:255-276): coefficient i moves to
// e = i g mod 2N, negated mod its prime when e >= N (X^N = -1). Written as a gather so every
// output word is stored once, coalesced: output j takes input i = j g^-1 mod 2N (g is odd, so
// invertible mod 2N), as it is when i < N and negated from i - N otherwise, since (i - N) g =
// j - N g = j + N mod 2N for odd g. Limb r of the layout is Q prime r for r <= level, else P prime
// nq + r - level - 1.
__global__ __launch_bounds__(256) void k_automorphism_coeff(const u64* __restrict__ in, u64* __restrict__ out,
                                                            u64 g_inv, u32 level, u32 nq,
                                                            const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const u32 r = blockIdx.y;
    const u64 q = pc[r <= level ? r : nq + r - level - 1].q;
    const u64 base = (u64)r * n;
    const u32 i = (u32)((j * g_inv) & (2 * (u64)n - 1));
    if (i < n) {
        out[base + j] = in[base + i];
    } else {
        const u64 c = in[base + i - n];
        out[base + j] = c == 0 ? 0 : q - c;
    }
}

void automorphism_coeff_launch(const u64* in, u64* out, u32 limbs, u32 level, u32 nq, u64 g, const PrimeConst* pc,
                               u32 log_n, hipStream_t stream) {
    if (limbs == 0) return;
    const u64 two_n = 2ull << log_n;
    u64 g_inv = 1;  // g^(2N/2 - 1) = g^-1 in the unit group mod 2N (order dividing N)
    for (u64 e = (two_n >> 1) - 1, b = g % two_n; e; e >>= 1, b = b * b % two_n)
        if (e & 1) g_inv = g_inv * b % two_n;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_automorphism_coeff, dim3((n + 255) / 256, limbs), dim3(256), 0, stream, in, out, g_inv, level,
                       nq, pc, log_n);
}

__global__ __launch_bounds__(256) void k_automorphism_batch(const AutJob* __restrict__ jobs, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const AutJob J = jobs[blockIdx.y];
    glb(J.out)[c] = glb(J.in)[galois_index(c, J.galois, log_n)];
}

void automorphism_batch_launch(const AutJob* jobs, u32 njobs, u32 log_n, hipStream_t stream) {
    if (njobs == 0) return;
    const u32 n = 1u << log_n;
    hipLaunchKernelGGL(k_automorphism_batch, dim3((n + 255) / 256, njobs), dim3(256), 0, stream, jobs, log_n);
}

// Stateless seeded generator: word (seed, stream, index) is uniform below q by rejection on
// a splitmix64-style mix of a counter (no modulo bias, as Rng::below in
// crates/ckks-train-core/src/sampler.rs:73-82).
CKH_DEV u64 mix64(u64 z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

CKH_DEV u64 uniform_word(u64 seed, u64 stream, u64 index, u64 q) {
    const u64 key = mix64(seed ^ mix64(stream + 0x9e3779b97f4a7c15ULL));
    const u64 zone = (~0ULL / q) * q;
    u64 x = 0;
    for (u64 a = 0; a < 64; ++a) {
        x = mix64(key + (index * 64 + a) * 0x9e3779b97f4a7c15ULL);
        if (x < zone) break;
    }
    return x % q;
}

__global__ __launch_bounds__(256) void k_fill_uniform(u64* __restrict__ out, u32 q_limbs, u32 nq, u64 seed,
                                                      u64 stream0, const PrimeConst* __restrict__ pc, u32 log_n) {
    const u32 n = 1u << log_n;
    const u32 c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const u32 li = blockIdx.y;
    const u32 prime = li < q_limbs ? li : nq + (li - q_limbs);
    out[(u64)li * n + c] = uniform_word(seed, stream0 + li, c, pc[prime].q);
}

void fill_uniform_launch(u64* out, u32 q_limbs, u32 p_limbs, u32 nq, u64 seed, u64 stream0, const PrimeConst* pc,
                         u32 log_n, hipStream_t stream) {
    const u32 n = 1u << log_n;
    if (q_limbs + p_limbs == 0) return;
    hipLaunchKernelGGL(k_fill_uniform, dim3((n + 255) / 256, q_limbs + p_limbs), dim3(256), 0, stream, out, q_limbs,
                       nq, seed, stream0, pc, log_n);
}

__global__ __launch_bounds__(256) void k_digest(const u64* __restrict__ data, u64 words, u64* out) {
    __shared__ u64 part[256];
    u64 acc = 0;
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < words; i += (u64)gridDim.x * 256)
        acc += mix64(data[i] ^ mix64(i + 0x632be59bd9b4e019ULL));
    part[threadIdx.x] = acc;
    __syncthreads();
    for (u32 w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)out, (unsigned long long)part[0]);
}

void digest_launch(const u64* data, u64 words, u64* out, hipStream_t stream) {
    u64 blocks = (words + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_digest, dim3((u32)blocks), dim3(256), 0, stream, data, words, out);
}

}  // namespace ckh
