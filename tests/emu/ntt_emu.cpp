// CPU replay of the register-radix NTT kernels (ckks-train_amd/csrc/ntt_passes.h, the code the
// gfx950 kernels run), one simulated workgroup of 256 threads at a time: the tile passes of a
// forward and an inverse transform of whole limbs, at N = 2^12 .. 2^16 (M = 6, 7, 8 for both
// passes), for 50-bit primes (the f64 arithmetic) and 60-bit primes (the integer arithmetic);
// then the fused row pass + key inner product + mod_down finish of the batched key switch
// (k_rows_ip_fin's body) over 1, 3, 4 and 7 digits, its sums and finish against the oracle.
//
// Checked (any failure aborts): every pass reads and writes every word of the limb exactly once;
// every LDS exchange writes and reads each slot once per phase, after it was written; every f64
// modular product and reduction meets its preconditions and equals the integer result; every
// Shoup product is below 2q and congruent; the words between the passes stay within the bounds
// the arithmetic relies on (f64 forward |x| <= 1.7 q after an even number of column stages, 2.6 q
// after an odd one, inverse |x| <= q; integer forward < 6q, inverse < 2q); the outputs equal the
// CPU oracle's transform (ckks_oracle.c, linked). An adversarial search on the column pass's
// inputs and a row pass fed words at 2.6 q (checked against the integer pass) probe those bounds.
// Test infrastructure only (built and run by tests/test_emu.py).
#include <algorithm>
#include <random>
#include <string>
#include <thread>

#include "emu_prelude.h"
#include "ntt_passes.h"
#include "ntt_tables.h"

extern "C" {
int or_ntt_forward_limbs(u64 n, const u64* q, u64 limbs, u64* a);
int or_ntt_inverse_limbs(u64 n, const u64* q, u64 limbs, u64* a);
int or_ntt_primes_skip(unsigned bits, u64 n, u64 skip, u64 count, u64* out);
}

using namespace ckh;
using namespace ckh::fast;

namespace {

// One workgroup: 256 threads running f, a barrier between them whose completion step checks the
// LDS phase that ends; the last phase is checked after the threads have finished.
template <class F>
void run_group(emu::Lds& lds, F f) {
    auto done = [&]() noexcept { lds.end_phase(); };
    std::barrier bar(256, done);
    emu::Group g{&lds, [&bar] { bar.arrive_and_wait(); }};
    emu::current() = &g;
    std::vector<std::thread> th;
    for (u32 t = 0; t < NT; ++t)
        th.emplace_back([&f, t] {
            threadIdx.x = t;
            f();
        });
    for (auto& x : th) x.join();
    lds.end_phase();
    emu::current() = nullptr;
}

void expect_once(emu::Tracked& t, bool reads, bool writes, const char* pass) {
    for (size_t i = 0; i < t.len; ++i) {
        const u32 r = t.reads[i], w = t.writes[i];
        if ((reads && r != 1) || (!reads && r != 0) || (writes && w != 1) || (!writes && w != 0)) {
            std::fprintf(stderr, "%s: %s word %zu read %u times, written %u times\n", pass, t.name, i, r, w);
            emu::die("a pass does not touch every word exactly once");
        }
    }
}

struct Limb {
    u32 log_n;
    u64 q;
    HostNttTables tab;
};

// Forward (or inverse) transform of one limb through the two passes, every tile replayed; yf
// (inverse only): the last stage leaves the y form of a conversion source (YForm).
void transform(const Limb& L, const std::vector<u64>& in, std::vector<u64>& out, bool forward, double* mid_max,
               const YForm* yf = nullptr) {
    const u32 log_n = L.log_n, B = log_n / 2, A = log_n - B, tiles = 1u << (log_n - TL);
    const u64 n = 1ull << log_n;
    std::vector<u64> src(in), dst(n, 0xdeadbeefdeadbeefull);
    emu::Tracked ts(src.data(), n, "src"), td(dst.data(), n, "dst");
    emu::registry() = {&ts, &td};
    emu::Lds lds(LDS_WORDS);
    const PrimeConst& pc = L.tab.pc[0];
    const bool fp = L.q < FP_PRIME_BOUND;
    const FpAr far{(forward ? L.tab.fwd_f : L.tab.inv_f).data(), pc.qf, pc.qinv, pc.n_inv_f, pc.w_n_inv_f,
                   (forward ? L.tab.rf : L.tab.ri).data()};
    const IntAr iar{(forward ? L.tab.fwd : L.tab.inv).data(), pc.q, pc.n_inv, pc.n_inv_s, pc.w_n_inv, pc.w_n_inv_s};
    auto pass = [&](u32 m, auto&& body, const char* name, bool rs, bool ws, bool rd, bool wd) {
        ts.reset();
        td.reset();
        for (u32 tile = 0; tile < tiles; ++tile) run_group(lds, [&] { body(m, tile); });
        expect_once(ts, rs, ws, name);
        expect_once(td, rd, wd, name);
    };
    emu::LdsView lv{&lds};
    auto dispatch = [&](u32 m, auto&& f) {
        switch (m) {
            case 6: f(std::integral_constant<u32, 6>{}); break;
            case 7: f(std::integral_constant<u32, 7>{}); break;
            default: f(std::integral_constant<u32, 8>{}); break;
        }
    };
    if (forward) {
        pass(A, [&](u32 m, u32 tile) {
            dispatch(m, [&](auto M) {
                if (fp) fwd_cols<decltype(M)::value>(far, src.data(), dst.data(), lv, log_n, tile);
                else fwd_cols<decltype(M)::value>(iar, src.data(), dst.data(), lv, log_n, tile);
            });
        }, "forward column pass", true, false, false, true);
    } else {
        pass(B, [&](u32 m, u32 tile) {
            dispatch(m, [&](auto M) {
                if (fp) inv_rows<decltype(M)::value>(far, src.data(), dst.data(), lv, log_n, tile);
                else inv_rows<decltype(M)::value>(iar, src.data(), dst.data(), lv, log_n, tile);
            });
        }, "inverse row pass", true, false, false, true);
    }
    // the words between the passes
    double mx = 0;
    for (u64 i = 0; i < n; ++i) {
        const double v = fp ? std::fabs(bits_f64(dst[i])) : (double)dst[i];
        if (fp && !emu::is_int(v)) emu::die("pass intermediate not an integer", v);
        mx = std::max(mx, v / (double)L.q);
    }
    // f64 forward: the column pass ends with a reduced stage (1.7 q) when it has an even number of
    // stages, with a lazy one (2.6 q) when odd
    const double bound = fp ? (forward ? (A % 2 ? 2.6 : 1.7) : 1.0) : (forward ? 6.0 : 2.0);
    if (fp ? mx > bound : mx >= bound) emu::die("pass intermediate above its bound", mx, bound);
    *mid_max = std::max(*mid_max, mx);
    if (forward) {
        pass(B, [&](u32 m, u32 tile) {
            dispatch(m, [&](auto M) {
                if (fp) fwd_rows<decltype(M)::value>(far, dst.data(), dst.data(), lv, log_n, nullptr, L.q, tile);
                else fwd_rows<decltype(M)::value>(iar, dst.data(), dst.data(), lv, log_n, nullptr, L.q, tile);
            });
        }, "forward row pass", false, false, true, true);
    } else {
        pass(A, [&](u32 m, u32 tile) {
            dispatch(m, [&](auto M) {
                constexpr u32 MM = decltype(M)::value;
                if (yf) {  // as k_inv_cols<M, true>
                    FpAr fy{far.tab, far.q, far.qinv, yf->n_f, yf->wn_f, far.rt};
                    fy.add = yf->h_f;
                    IntAr iy{iar.tab, iar.q, yf->n, yf->n_s, yf->wn, yf->wn_s};
                    iy.add = yf->h;
                    if (fp) inv_cols<MM, FpAr, false, true>(fy, dst.data(), lv, log_n, tile);
                    else inv_cols<MM, IntAr, false, true>(iy, dst.data(), lv, log_n, tile);
                } else {
                    if (fp) inv_cols<MM>(far, dst.data(), lv, log_n, tile);
                    else inv_cols<MM>(iar, dst.data(), lv, log_n, tile);
                }
            });
        }, "inverse column pass", false, false, true, true);
    }
    emu::registry().clear();
    out = dst;
}

// The column pass alone (forward): the row pass's input, as the device leaves it in memory.
std::vector<u64> cols_only(const Limb& L, const std::vector<u64>& in) {
    const u32 log_n = L.log_n, A = log_n - log_n / 2, tiles = 1u << (log_n - TL);
    const u64 n = 1ull << log_n;
    std::vector<u64> src(in), dst(n, 0);
    emu::Tracked ts(src.data(), n, "src"), td(dst.data(), n, "dst");
    emu::registry() = {&ts, &td};
    emu::Lds lds(LDS_WORDS);
    emu::LdsView lv{&lds};
    const PrimeConst& pc = L.tab.pc[0];
    const FpAr far{L.tab.fwd_f.data(), pc.qf, pc.qinv, pc.n_inv_f, pc.w_n_inv_f, L.tab.rf.data()};
    for (u32 tile = 0; tile < tiles; ++tile)
        run_group(lds, [&] {
            switch (A) {
                case 6: fwd_cols<6>(far, src.data(), dst.data(), lv, log_n, tile); break;
                case 7: fwd_cols<7>(far, src.data(), dst.data(), lv, log_n, tile); break;
                default: fwd_cols<8>(far, src.data(), dst.data(), lv, log_n, tile); break;
            }
        });
    emu::registry().clear();
    return dst;
}

// The forward row pass alone, in place, with the f64 or the integer arithmetic.
std::vector<u64> rows_only(const Limb& L, std::vector<u64> mid, bool fp) {
    const u32 log_n = L.log_n, M = log_n / 2, tiles = 1u << (log_n - TL);
    const u64 n = 1ull << log_n;
    emu::Tracked tm(mid.data(), n, "mid");
    emu::registry() = {&tm};
    emu::Lds lds(LDS_WORDS);
    emu::LdsView lv{&lds};
    const PrimeConst& pc = L.tab.pc[0];
    const FpAr far{L.tab.fwd_f.data(), pc.qf, pc.qinv, pc.n_inv_f, pc.w_n_inv_f, L.tab.rf.data()};
    const IntAr iar{L.tab.fwd.data(), pc.q, pc.n_inv, pc.n_inv_s, pc.w_n_inv, pc.w_n_inv_s};
    for (u32 tile = 0; tile < tiles; ++tile)
        run_group(lds, [&] {
            auto body = [&](auto Mc) {
                constexpr u32 MM = decltype(Mc)::value;
                if (fp) fwd_rows<MM>(far, mid.data(), mid.data(), lv, log_n, nullptr, L.q, tile);
                else fwd_rows<MM>(iar, mid.data(), mid.data(), lv, log_n, nullptr, L.q, tile);
            };
            switch (M) {
                case 6: body(std::integral_constant<u32, 6>{}); break;
                case 7: body(std::integral_constant<u32, 7>{}); break;
                default: body(std::integral_constant<u32, 8>{}); break;
            }
        });
    emu::registry().clear();
    return mid;
}

// Adversarial inputs for the f64 forward transform of a 50-bit prime at N = 2^log_n.
// Column pass: coordinate ascent on the inputs of every column at once (columns are independent
// in the column pass), a change kept where it raises that column's largest word between the
// passes; the largest found must stay within the bound the arithmetic relies on (1.7 q after an
// even number of column stages, 2.6 q after an odd one). Row pass: words at and near 2.6 q with
// random signs; every product and reduction is
// checked exact by the prelude's hooks, and the outputs must equal the integer-arithmetic pass
// (IntAr, the 60-bit primes' Harvey butterflies) run on the same words reduced mod q.
void adversarial_f64(std::mt19937_64& rng, u32 log_n, int iters) {
    const u32 B = log_n / 2, A = log_n - B;
    const u64 n = 1ull << log_n, cols = 1ull << B, rows = n / cols;
    u64 q = 0;
    if (or_ntt_primes_skip(50, n, 0, 1, &q) != 0) emu::die("no prime");
    Limb L{log_n, q, build_ntt_tables({q}, log_n, true)};
    auto col_max = [&](const std::vector<u64>& mid, std::vector<double>& m) {
        m.assign(cols, 0.0);
        for (u64 r = 0; r < rows; ++r)
            for (u64 c = 0; c < cols; ++c) m[c] = std::max(m[c], std::fabs(bits_f64(mid[r * cols + c])) / (double)q);
    };
    const u64 pool[] = {0, q - 1, (q - 1) / 2, (q + 1) / 2, 1, q - 2};
    std::vector<u64> x(n);
    for (auto& v : x) v = pool[rng() % 6];
    std::vector<double> best, cur;
    col_max(cols_only(L, x), best);
    for (int it = 0; it < iters; ++it) {
        std::vector<u64> y(x);
        for (u64 c = 0; c < cols; ++c)
            for (int k = 0; k < 2; ++k) {
                const u64 r = rng() % rows;
                y[r * cols + c] = (rng() & 3) ? pool[rng() % 6] : rng() % q;
            }
        col_max(cols_only(L, y), cur);
        for (u64 c = 0; c < cols; ++c)
            if (cur[c] > best[c]) {
                best[c] = cur[c];
                for (u64 r = 0; r < rows; ++r) x[r * cols + c] = y[r * cols + c];
            }
    }
    const double found = *std::max_element(best.begin(), best.end()), bound = A % 2 ? 2.6 : 1.7;
    if (found > bound) emu::die("column pass above its bound on adversarial input", found, bound);
    // the row pass at the larger of the two between-pass bounds (its first stage is lazy: up to
    // 3.7 q inside it)
    const double lim = std::floor(2.6 * (double)q);
    double row_in_max = 0;
    for (int rep = 0; rep < 4; ++rep) {
        std::vector<u64> mid(n), red(n);
        for (u64 i = 0; i < n; ++i) {
            double v = (rng() & 1) ? lim : lim - (double)(rng() % 1024);
            if (rep == 1) v = (double)(rng() % (u64)lim);
            if (rng() & 1) v = -v;
            mid[i] = f64_bits(v);
            const long long r = (long long)v % (long long)q;
            red[i] = (u64)(r < 0 ? r + (long long)q : r);
            row_in_max = std::max(row_in_max, std::fabs(v) / (double)q);
        }
        const std::vector<u64> f = rows_only(L, mid, true), g = rows_only(L, red, false);
        for (u64 i = 0; i < n; ++i)
            if (f[i] >= q || f[i] != g[i]) emu::die("f64 row pass at the bound differs from the integer pass", (double)i);
        // the integer row pass at its input bound (below 6 q: the same residues plus 5 q or 4 q)
        std::vector<u64> big(n);
        for (u64 i = 0; i < n; ++i) big[i] = red[i] + ((rng() & 3) ? 5 : 4) * q;
        const std::vector<u64> h = rows_only(L, big, false);
        for (u64 i = 0; i < n; ++i)
            if (h[i] != g[i]) emu::die("integer row pass on words below 6 q differs", (double)i);
    }
    std::printf("adversarial f64 inputs at N=2^%u: column pass max |x|/q %.3f after %d rounds of ascent (bound "
                "%.1f); row pass on words up to %.3f q equals the integer pass, as does that on words below 6 q\n",
                log_n, found, iters, bound, row_in_max);
    std::fflush(stdout);
}

std::vector<u64> oracle(const Limb& L, std::vector<u64> a, bool forward);

// The forward row pass with the finish of mod_down and rescale (k_fwd_rows<M, true>, f64 path):
// x's transform (from its column pass) against a and an optional addend, out = (a - NTT(x)) inv
// (+ add). Pattern 0: random words, 1: all q - 1, 2: random without an addend. Every word of a,
// the addend and the column-pass output is read once and every output word written once; the
// outputs equal the same computed with the oracle's transform.
void replay_finish(const Limb& L, std::mt19937_64& rng, int pattern) {
    const u32 log_n = L.log_n, M = log_n / 2, tiles = 1u << (log_n - TL);
    const u64 n = 1ull << log_n, q = L.q;
    std::vector<u64> x(n), a(n), add(n), out(n, 0x5a5a5a5a5a5a5a5aull);
    for (auto* v : {&x, &a, &add})
        for (auto& w : *v) w = pattern == 1 ? q - 1 : rng() % q;
    std::vector<u64> mid = cols_only(L, x);
    const u64 inv = 1 + rng() % (q - 1);
    FinishJob F{a.data(), nullptr, out.data(), pattern == 2 ? nullptr : add.data(), inv,
                (u64)(((unsigned __int128)inv << 64) / q), 0, 0, {(double)inv, (double)inv / (double)q}};
    emu::Tracked tm(mid.data(), n, "mid"), ta(a.data(), n, "a"), tad(add.data(), n, "addend"), to(out.data(), n, "out");
    emu::registry() = {&tm, &ta, &tad, &to};
    emu::Lds lds(LDS_WORDS);
    emu::LdsView lv{&lds};
    const PrimeConst& pc = L.tab.pc[0];
    const FpAr far{L.tab.fwd_f.data(), pc.qf, pc.qinv, pc.n_inv_f, pc.w_n_inv_f, L.tab.rf.data()};
    for (u32 tile = 0; tile < tiles; ++tile)
        run_group(lds, [&] {
            switch (M) {
                case 6: fwd_rows<6>(far, mid.data(), mid.data(), lv, log_n, &F, q, tile); break;
                case 7: fwd_rows<7>(far, mid.data(), mid.data(), lv, log_n, &F, q, tile); break;
                default: fwd_rows<8>(far, mid.data(), mid.data(), lv, log_n, &F, q, tile); break;
            }
        });
    emu::registry().clear();
    for (u64 i = 0; i < n; ++i) {
        if (tm.reads[i] != 1 || tm.writes[i] != 0) emu::die("finish: a transform word not read exactly once", (double)i);
        if (ta.reads[i] != 1 || ta.writes[i] != 0) emu::die("finish: an a word not read exactly once", (double)i);
        if (tad.reads[i] != (pattern == 2 ? 0u : 1u) || tad.writes[i] != 0)
            emu::die("finish: an addend word not read exactly once", (double)i);
        if (to.writes[i] != 1 || to.reads[i] != 0) emu::die("finish: an output word not written exactly once", (double)i);
    }
    const std::vector<u64> t = oracle(L, x, true);
    for (u64 i = 0; i < n; ++i) {
        u64 w = (u64)((unsigned __int128)((a[i] + q - t[i]) % q) * inv % q);
        if (pattern != 2) w = (w + add[i]) % q;
        if (out[i] != w) emu::die("finish output differs from the oracle", log_n, (double)i);
    }
}

// The fused row pass + inner product + mod_down finish of one limb (k_rows_ip_fin's body): ndig
// digits (digit own_j given in NTT form, the others as column-pass outputs), key parts, the
// converted limbs z_b / z_a as column-pass outputs, an addend for b (and, with pattern 1, for a). Every word of every
// input is read exactly once and every output word written once; the outputs equal
// ((sum_j X_j k_j) - NTT(z)) P^-1 (+ add) computed with the oracle's transforms.
void replay_rows_ip_fin(const Limb& L, u32 ndig, u32 own_j, std::mt19937_64& rng, int pattern) {
    const u32 log_n = L.log_n, M = log_n / 2, tiles = 1u << (log_n - TL);
    const u64 n = 1ull << log_n, q = L.q;
    auto rnd = [&](std::vector<u64>& v) {
        for (auto& x : v) x = pattern == 1 ? q - 1 : rng() % q;
    };
    std::vector<u64> mid((size_t)ndig * n), own(n), key((size_t)2 * ndig * n), zb_mid, za_mid, add(n), add2(n),
        outb(n, 7), outa(n, 7);
    std::vector<std::vector<u64>> X(ndig);
    for (u32 j = 0; j < ndig; ++j) {
        std::vector<u64> x(n);
        rnd(x);
        if (j == own_j) {
            std::copy(x.begin(), x.end(), own.begin());
            X[j] = x;
        } else {
            const std::vector<u64> m = cols_only(L, x);
            std::copy(m.begin(), m.end(), mid.begin() + (size_t)j * n);
            X[j] = oracle(L, x, true);
        }
    }
    rnd(key);
    std::vector<u64> zb(n), za(n);
    rnd(zb);
    rnd(za);
    zb_mid = cols_only(L, zb);
    za_mid = cols_only(L, za);
    rnd(add);
    rnd(add2);
    const bool both = pattern == 1;  // an addend on both sums (relinearization), else on b only (rotation)
    const u64 inv = 1 + rng() % (q - 1);
    RowIpFinJob J{};
    J.ip = {mid.data(), n, own.data(), key.data(), n, nullptr, nullptr, ndig, 0, own_j, 0};
    J.zb = zb_mid.data();
    J.za = za_mid.data();
    J.out_b = outb.data();
    J.out_a = outa.data();
    J.add_b = add.data();
    J.add_a = both ? add2.data() : nullptr;
    J.inv = {(double)inv, (double)inv / (double)q};
    emu::Tracked tm(mid.data(), mid.size(), "digits"), to(own.data(), n, "own digit"), tk(key.data(), key.size(), "key"),
        tzb(zb_mid.data(), n, "z_b"), tza(za_mid.data(), n, "z_a"), tad(add.data(), n, "addend"),
        tad2(add2.data(), n, "addend a"), tob(outb.data(), n, "out_b"), toa(outa.data(), n, "out_a");
    emu::registry() = {&tm, &to, &tk, &tzb, &tza, &tad, &tad2, &tob, &toa};
    emu::Lds lds(LDS_WORDS);
    emu::LdsView lv{&lds};
    const PrimeConst& pc = L.tab.pc[0];
    const FpAr far{L.tab.fwd_f.data(), pc.qf, pc.qinv, pc.n_inv_f, pc.w_n_inv_f, L.tab.rf.data()};
    for (u32 tile = 0; tile < tiles; ++tile)
        run_group(lds, [&] {
            auto body = [&](auto Mc) {
                constexpr u32 MM = decltype(Mc)::value;
                double sb[16], sa[16];
                rows_ip_sums<MM>(far, J.ip, lv, log_n, tile, sb, sa);
                finish_sum<MM>(far, sb, J.zb, J.out_b, J.add_b, J.inv, lv, log_n, tile);
                finish_sum<MM>(far, sa, J.za, J.out_a, J.add_a, J.inv, lv, log_n, tile);
            };
            switch (M) {
                case 6: body(std::integral_constant<u32, 6>{}); break;
                case 7: body(std::integral_constant<u32, 7>{}); break;
                default: body(std::integral_constant<u32, 8>{}); break;
            }
        });
    emu::registry().clear();
    // coverage: the own digit's and the other digits' words read once (digit own_j's slot of mid
    // is not read), the key, z and the addend read once, the outputs written once
    for (u32 j = 0; j < ndig; ++j)
        for (u64 i = 0; i < n; ++i) {
            const u32 want = j == own_j ? 0 : 1;
            if (tm.reads[(size_t)j * n + i] != want || tm.writes[(size_t)j * n + i] != 0)
                emu::die("a digit word not read exactly once", j, (double)i);
        }
    for (emu::Tracked* t : {&to, &tk, &tzb, &tza, &tad})
        for (size_t i = 0; i < t->len; ++i)
            if (t->reads[i] != 1 || t->writes[i] != 0) emu::die("an input word not read exactly once", (double)i);
    for (size_t i = 0; i < n; ++i)
        if (tad2.reads[i] != (both ? 1u : 0u) || tad2.writes[i] != 0) emu::die("an addend word not read once", (double)i);
    for (emu::Tracked* t : {&tob, &toa})
        for (size_t i = 0; i < t->len; ++i)
            if (t->writes[i] != 1 || t->reads[i] != 0) emu::die("an output word not written exactly once", (double)i);
    const std::vector<u64> Zb = oracle(L, zb, true), Za = oracle(L, za, true);
    auto mulm = [&](u64 a, u64 b) { return (u64)((unsigned __int128)a * b % q); };
    for (u64 i = 0; i < n; ++i) {
        u64 sbv = 0, sav = 0;
        for (u32 j = 0; j < ndig; ++j) {
            sbv = (sbv + mulm(X[j][i], key[(size_t)2 * j * n + i])) % q;
            sav = (sav + mulm(X[j][i], key[(size_t)(2 * j + 1) * n + i])) % q;
        }
        const u64 wb = (mulm((sbv + q - Zb[i]) % q, inv) + add[i]) % q,
                  wa = (mulm((sav + q - Za[i]) % q, inv) + (both ? add2[i] : 0)) % q;
        if (outb[i] != wb || outa[i] != wa) emu::die("fused row pass + inner product + finish differs", (double)i);
    }
}

std::vector<u64> oracle(const Limb& L, std::vector<u64> a, bool forward) {
    const u64 n = 1ull << L.log_n;
    if ((forward ? or_ntt_forward_limbs : or_ntt_inverse_limbs)(n, &L.q, 1, a.data()) != 0)
        emu::die("oracle transform failed");
    return a;
}

}  // namespace

int main(int argc, char** argv) {
    // argv[1]: random inputs per (N, prime); argv[2]: random +-(q-1)/2 sign patterns at N = 2^12;
    // argv[3]: a planted mistake (0 none, 1 colliding LDS slots, 2 colliding global words, 3 one
    // wrong row factor, 4 a row factor whose w and w/q disagree), run at N = 2^12, 50-bit primes
    const int reps = argc > 1 ? std::atoi(argv[1]) : 1;
    const int patterns = argc > 2 ? std::atoi(argv[2]) : 8;
    const int plant = argc > 3 ? std::atoi(argv[3]) : 0;
    if (plant == 1 || plant == 2) emu::mutation() = plant;
    std::mt19937_64 rng(20261203);
    for (u32 log_n = 12; log_n <= (plant ? 12u : 16u); ++log_n) {
        const u64 n = 1ull << log_n;
        for (unsigned bits : {50u, 60u}) {
            if (plant && bits != 50) continue;
            u64 q = 0;
            if (or_ntt_primes_skip(bits, n, 0, 1, &q) != 0) emu::die("no prime");
            Limb L{log_n, q, build_ntt_tables({q}, log_n, true)};
            if (plant == 3) L.tab.rf[5 * ROW_FACTOR_SLOTS + 3] = L.tab.rf[6 * ROW_FACTOR_SLOTS + 3];
            if (plant == 4) L.tab.rf[5 * ROW_FACTOR_SLOTS + 3].w = L.tab.rf[6 * ROW_FACTOR_SLOTS + 3].w;
            std::vector<std::vector<u64>> inputs;
            for (int r = 0; r < reps; ++r) {
                std::vector<u64> a(n);
                for (auto& x : a) x = rng() % q;
                inputs.push_back(a);
            }
            inputs.push_back(std::vector<u64>(n, q - 1));
            inputs.push_back(std::vector<u64>(n, 0));
            {
                std::vector<u64> a(n);
                for (u64 i = 0; i < n; ++i) a[i] = (i & 1) ? (q - 1) / 2 : (q + 1) / 2;
                inputs.push_back(a);
            }
            if (log_n == 12)
                for (int p = 0; p < patterns; ++p) {
                    std::vector<u64> a(n);
                    for (auto& x : a) x = (rng() & 1) ? (q - 1) / 2 : (q + 1) / 2;
                    inputs.push_back(a);
                }
            // the y form of a conversion source with a random (Q/q)^-1 (h = q >> 1 as in every plan)
            YForm yf{};
            const u64 qhinv = 1 + rng() % (q - 1), h = q >> 1;
            {
                const PrimeConst& pc = L.tab.pc[0];
                auto mm = [q](u64 a, u64 b) { return (u64)((unsigned __int128)a * b % q); };
                auto sh = [q](u64 w) { return (u64)(((unsigned __int128)w << 64) / q); };
                yf.n = mm(pc.n_inv, qhinv);
                yf.wn = mm(pc.w_n_inv, qhinv);
                yf.h = mm(h, qhinv);
                yf.n_s = sh(yf.n);
                yf.wn_s = sh(yf.wn);
                yf.n_f = {(double)yf.n, (double)yf.n / (double)q};
                yf.wn_f = {(double)yf.wn, (double)yf.wn / (double)q};
                yf.h_f = (double)yf.h;
            }
            double fmax = 0, imax = 0;
            for (const auto& a : inputs) {
                std::vector<u64> f, b, y;
                transform(L, a, f, true, &fmax);
                if (f != oracle(L, a, true)) emu::die("forward output differs from the oracle", log_n, bits);
                transform(L, f, b, false, &imax);
                if (b != a) emu::die("inverse does not return the input", log_n, bits);
                transform(L, f, y, false, &imax, &yf);
                for (u64 i = 0; i < n; ++i)
                    if (y[i] != (u64)((unsigned __int128)((a[i] + h) % q) * qhinv % q))
                        emu::die("y-form inverse output differs from [x + h] (Q/q)^-1", log_n, bits, (double)i);
            }
            std::printf("N=2^%u q=%u-bit: %zu limbs forward and inverse exact, inverse in y form exact; between "
                        "passes max |x|/q forward %.3f inverse %.3f\n",
                        log_n, bits, inputs.size(), fmax, imax);
            std::fflush(stdout);
        }
    }
    if (!plant) {
        for (u32 log_n = 12; log_n <= 16; ++log_n) {
            u64 q = 0;
            if (or_ntt_primes_skip(50, 1ull << log_n, 0, 1, &q) != 0) emu::die("no prime");
            Limb L{log_n, q, build_ntt_tables({q}, log_n, true)};
            for (int pattern = 0; pattern < 3; ++pattern) replay_finish(L, rng, pattern);
            std::printf("N=2^%u: row pass + finish (a - NTT(x)) inv (+ add) exact (random, all q - 1, no addend)\n",
                        log_n);
            std::fflush(stdout);
        }
        for (u32 log_n : {12u, 13u}) adversarial_f64(rng, log_n, 200);
    }
    if (!plant) {
        // the fused row pass + inner product + finish, 50-bit primes, N = 2^12 .. 2^16
        for (u32 log_n = 12; log_n <= 16; ++log_n) {
            const u64 n = 1ull << log_n;
            u64 q = 0;
            if (or_ntt_primes_skip(50, n, 0, 1, &q) != 0) emu::die("no prime");
            Limb L{log_n, q, build_ntt_tables({q}, log_n, true)};
            for (auto [ndig, pattern] : {std::pair<u32, int>{1, 0}, {3, 0}, {4, 0}, {4, 1}, {7, 1}})
                replay_rows_ip_fin(L, ndig, ndig / 2, rng, pattern);
            std::printf("N=2^%u: fused row pass + inner product + finish exact (1, 3, 4, 7 digits; all q - 1)\n",
                        log_n);
            std::fflush(stdout);
        }
    }
    auto& s = emu::stats();
    std::printf("f64 products %llu (max |y|/q %.3f, max |t|/q %.3f), reductions %llu (max |x|/q %.3f), "
                "Shoup products %llu\nEMU OK\n",
                (unsigned long long)s.fmul_calls, s.fmul_in, s.fmul_out, (unsigned long long)s.fred_calls, s.fred_in,
                (unsigned long long)s.shoup_calls);
    return 0;
}
