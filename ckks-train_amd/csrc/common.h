// Device prelude of the gfx950 RNS-CKKS library: the HIP runtime, the device forms of the few
// primitives the portable arithmetic (arith.h: descriptors, per-prime constants, modular
// arithmetic) leaves to its includer, and device-only helpers (address-space casts, XCD tiling).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

namespace ckh {
#define CKH_DEV __device__ __forceinline__
// High word of a 64x64 product (the compiler builds it from 32-bit multiplies on gfx950).
CKH_DEV u64 mulhi(u64 a, u64 b) { return __umul64hi(a, b); }
// The bits of a u64 as an f64 and back (signed pass intermediates are kept as raw f64 bits).
CKH_DEV double bits_f64(u64 x) { return __longlong_as_double((long long)x); }
CKH_DEV u64 f64_bits(double d) { return (u64)__double_as_longlong(d); }
// 32-bit bit reversal.
CKH_DEV u32 brev32(u32 x) { return __brev(x); }
}  // namespace ckh
// A value the optimizer must treat as unknown (it cannot hoist what is derived from it).
#define CKH_OPAQUE(x) asm volatile("" : "+v"(x))

#include "arith.h"

namespace ckh {


// A pointer to memory that no thread writes during the kernel (tables, plans), as a constant
// address-space pointer: loads through it may use the scalar cache even after the kernel's own
// stores, which the compiler otherwise cannot prove do not alias.
template <class T>
CKH_DEV const __attribute__((address_space(4))) T* cst(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

// The integer-reduction constants of a prime, read through the scalar cache.
CKH_DEV PrimeConst pc_int(const PrimeConst* p) {
    const auto* c = cst(p);
    PrimeConst r{};
    r.q = c->q;
    r.one_s = c->one_s;
    r.two64 = c->two64;
    r.two64_s = c->two64_s;
    return r;
}

// A device pointer read from a descriptor, as a global address-space pointer: accesses through
// it compile to global (not flat) memory instructions, which do not share the scalar loads'
// wait counter.
template <class T>
CKH_DEV __attribute__((address_space(1))) T* glb(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
// A streaming (non-temporal) store through a glb() view: word i of p.
template <class T>
CKH_DEV void st_nt(__attribute__((address_space(1))) T* p, u64 i, T v) {
    __builtin_nontemporal_store(v, p + i);
}

// blockIdx.x remapped so that consecutive logical tiles of one launch land on one XCD
// (workgroups are dealt round-robin over the 8 XCDs, each with its own L2).
CKH_DEV u32 xcd_tile(u32 bid, u32 nblocks) {
    const u32 nx = 8;
    if (nblocks % nx != 0) return bid;
    return (bid % nx) * (nblocks / nx) + bid / nx;
}

}  // namespace ckh
