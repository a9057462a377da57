// Host prelude for the CPU replay of the device kernels: the register-radix NTT passes and the
// vector base conversion (test infrastructure).
//
// It gives the kernel sources (ckks-train_amd/csrc/arith.h, ntt_passes.h and conv_arith.h,
// included unchanged) what the device prelude gives them on the GPU, with checks added:
//  * a workgroup is 256 host threads; __syncthreads() is a barrier across them, at which the
//    LDS accesses of the phase that ends are checked (every slot written at most once, read at
//    most once, and read only if written in the phase before);
//  * glb(p) returns a view of a registered array that counts reads and writes per element, so a
//    whole pass can be checked to read and write every element exactly once;
//  * the f64 modular product, the f64 reduction and the 64-bit Shoup product check their
//    preconditions and verify every result exactly against 128-bit integer arithmetic.
// Any violated check aborts with a message (the calling test fails).
#pragma once

#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <vector>

typedef uint64_t u64;
typedef uint32_t u32;

namespace emu {

[[noreturn]] inline void die(const char* what, double a = 0, double b = 0, double c = 0) {
    std::fprintf(stderr, "EMU CHECK FAILED: %s (%.17g, %.17g, %.17g)\n", what, a, b, c);
    std::fflush(stderr);
    std::abort();
}

// Largest magnitudes seen, in units of q (updated by the arithmetic checks).
struct Stats {
    std::mutex mu;
    double fmul_in = 0, fmul_out = 0, fred_in = 0, fmul_in_abs = 0;
    u64 fmul_calls = 0, fred_calls = 0, shoup_calls = 0;
    void note(double& slot, double v) {
        std::lock_guard<std::mutex> lk(mu);
        if (v > slot) slot = v;
    }
};
inline Stats& stats() {
    static Stats s;
    return s;
}

inline bool is_int(double x) { return x == std::rint(x); }

// A planted mistake for the replay's own tests (0: none): 1 = LDS slots 2k and 2k + 1 collide
// (an index map that is not one-to-one), 2 = global words 2k and 2k + 1 collide.
inline int& mutation() {
    static int m = 0;
    return m;
}

inline __int128 i128(double x) { return (__int128)(long long)x; }

inline void check_fmulmod(double y, double w, double wq, double q, double t) {
    if (!is_int(y) || !is_int(w) || !is_int(t)) die("fmulmod operand or result not an integer", y, w, t);
    // |y| < 2^53 keeps y an exact integer; the exactness of t is checked below (an inexact
    // step would leave t incongruent or outside the bound)
    if (std::fabs(y) >= 0x1p53) die("fmulmod |y| not below 2^53", y, q);
    if (std::fabs(w) >= q) die("fmulmod |w| not below q", w, q);
    // t = y w - k q exactly, with |t| <= (1/2 + 3 |y| / (16 q)) q (+1 for the rounding of k)
    const __int128 diff = i128(y) * i128(w) - i128(t);
    const __int128 qq = (__int128)(long long)q;
    if (diff % qq != 0) die("fmulmod result not congruent to y w", y, w, t);
    const double bound = (0.5 + 3.0 * std::fabs(y) / (16.0 * q)) * q + 1.0;
    if (std::fabs(t) > bound) die("fmulmod result above its bound", y, t, bound);
    auto& s = stats();
    s.note(s.fmul_in, std::fabs(y) / q);
    s.note(s.fmul_out, std::fabs(t) / q);
    std::lock_guard<std::mutex> lk(s.mu);
    ++s.fmul_calls;
}

inline void check_freduce(double x, double q, double r) {
    if (!is_int(x) || !is_int(r)) die("freduce operand or result not an integer", x, r);
    // x an integer below 2^53: RN(x RN(1/q)) misses x / q by at most 2 |x| 2^-53 / q < 2 / q,
    // so k = rint(.) is within 1/2 + 2/q of x / q and |x - k q| <= q/2 + 2 (computed exactly)
    if (std::fabs(x) >= 0x1p53) die("freduce |x| not below 2^53", x, q);
    if ((i128(x) - i128(r)) % (__int128)(long long)q != 0) die("freduce result not congruent", x, r, q);
    if (std::fabs(r) > 0.5 * q + 2.0) die("freduce result above q/2", x, r, q);
    auto& s = stats();
    s.note(s.fred_in, std::fabs(x) / q);
    std::lock_guard<std::mutex> lk(s.mu);
    ++s.fred_calls;
}

inline void check_shoup(u64 a, u64 w, u64 ws, u64 q, u64 r) {
    if (w >= q) die("shoup twiddle not reduced", (double)w, (double)q);
    if (ws != (u64)(((unsigned __int128)w << 64) / q)) die("shoup companion wrong", (double)w, (double)ws);
    if (r >= 2 * q) die("shoup product not below 2q", (double)a, (double)r, (double)q);
    if (((unsigned __int128)a * w) % q != r % q) die("shoup product not congruent", (double)a, (double)w, (double)r);
    auto& s = stats();
    std::lock_guard<std::mutex> lk(s.mu);
    ++s.shoup_calls;
}

}  // namespace emu

#define CKH_CHECK_FMULMOD(y, w, wq, q, t) emu::check_fmulmod(y, w, wq, q, t)
#define CKH_CHECK_FREDUCE(x, q, r) emu::check_freduce(x, q, r)
#define CKH_CHECK_SHOUP(a, w, ws, q, r) emu::check_shoup(a, w, ws, q, r)
#define CKH_OPAQUE(x) ((void)(x))

namespace ckh {
#define CKH_DEV inline
inline u64 mulhi(u64 a, u64 b) { return (u64)(((unsigned __int128)a * b) >> 64); }
inline double bits_f64(u64 x) {
    double d;
    std::memcpy(&d, &x, 8);
    return d;
}
inline u64 f64_bits(double d) {
    u64 x;
    std::memcpy(&x, &d, 8);
    return x;
}
inline u32 brev32(u32 x) {
    u32 r = 0;
    for (u32 i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
}  // namespace ckh

#include "arith.h"

namespace emu {

// A registered array: per-element read and write counts of the current pass.
struct Tracked {
    const u64* base;
    size_t len;
    const char* name;
    std::vector<std::atomic<u32>> reads, writes;
    Tracked(const u64* b, size_t n, const char* nm) : base(b), len(n), name(nm), reads(n), writes(n) {}
    void reset() {
        for (auto& c : reads) c = 0;
        for (auto& c : writes) c = 0;
    }
};

inline std::vector<Tracked*>& registry() {
    static std::vector<Tracked*> r;
    return r;
}

inline Tracked* find(const u64* p) {
    for (Tracked* t : registry())
        if (p >= t->base && p < t->base + t->len) return t;
    die("access outside every registered array");
}

// One element of a tracked array: reading it counts a read, assigning to it a write.
struct Ref {
    u64* p;
    Tracked* t;
    size_t i;
    operator u64() const {
        ++t->reads[i];
        return *p;
    }
    Ref& operator=(u64 v) {
        ++t->writes[i];
        *p = v;
        return *this;
    }
};

template <class T>
struct View {
    T* p;
    Tracked* t;
    Ref operator[](u64 i) const {
        if (!t) die("access through a null pointer", (double)i);
        u64* e = const_cast<u64*>(reinterpret_cast<const u64*>(p)) + i;
        if (mutation() == 2) e = const_cast<u64*>(t->base) + ((size_t)(e - t->base) & ~(size_t)1);
        const size_t at = (size_t)(e - t->base);
        if (e < t->base || at >= t->len) die("index outside the registered array", (double)i);
        return Ref{e, t, at};
    }
};

// The workgroup's LDS: counts per slot and phase, checked at every barrier.
struct Lds {
    std::vector<u64> w;
    std::vector<std::atomic<u32>> rd, wr;
    std::vector<u32> prev_wr;  // slots written in the previous phase
    explicit Lds(size_t n) : w(n), rd(n), wr(n), prev_wr(n) {}
    // Called once per barrier (by one thread, all others waiting).
    void end_phase() {
        for (size_t i = 0; i < w.size(); ++i) {
            const u32 r = rd[i], x = wr[i];
            if (x > 1) die("LDS slot written twice in one phase", (double)i);
            if (r > 1) die("LDS slot read twice in one phase", (double)i);
            if (r && x) die("LDS slot read and written in one phase", (double)i);
            if (r && !prev_wr[i]) die("LDS slot read without a write in the phase before", (double)i);
        }
        for (size_t i = 0; i < w.size(); ++i) {
            prev_wr[i] = wr[i];
            rd[i] = 0;
            wr[i] = 0;
        }
    }
};

struct LdsRef {
    Lds* l;
    size_t i;
    operator u64() const {
        ++l->rd[i];
        return l->w[i];
    }
    LdsRef& operator=(u64 v) {
        ++l->wr[i];
        l->w[i] = v;
        return *this;
    }
};

struct LdsView {
    Lds* l;
    LdsRef operator[](u32 i) const {
        if (mutation() == 1) i &= ~1u;
        if (i >= l->w.size()) die("LDS index out of range", (double)i);
        return LdsRef{l, i};
    }
};

// The workgroup being replayed: its barrier (with the LDS check as completion step).
struct Group {
    Lds* lds;
    std::function<void()> sync;
};
inline Group*& current() {
    static Group* g = nullptr;
    return g;
}

}  // namespace emu

struct EmuThreadIdx {
    u32 x;
};
inline thread_local EmuThreadIdx threadIdx{0};

inline void __syncthreads() { emu::current()->sync(); }

namespace ckh {
template <class T>
emu::View<T> glb(T* p) {
    // a view of a null pointer (an absent optional operand) is allowed; accessing it is not
    return emu::View<T>{p, p ? emu::find(reinterpret_cast<const u64*>(p)) : nullptr};
}
template <class T>
void st_nt(emu::View<T> p, u64 i, T v) {
    p[i] = v;
}
namespace fast {
inline u64 load_coherent(const u64* p) { return glb(p)[0]; }
}  // namespace fast
// Read-only tables: plain host pointers (the scalar cache on the device).
template <class T>
const T* cst(const T* p) {
    return p;
}
inline PrimeConst pc_int(const PrimeConst* p) { return *p; }
}  // namespace ckh
