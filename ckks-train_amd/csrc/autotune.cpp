// The project's autotuner: picks launch schedules per context by timing candidates on the
// device (HIP events on the context's stream) with synthetic limbs of the context's primes.
#include <algorithm>
#include <sstream>

#include "evaluator.h"

namespace ckh {

// Milliseconds per call of f over reps calls queued back to back on s (HIP events).
template <class F>
static double time_on(hipStream_t s, u32 reps, F&& f) {
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    hip_check(hipEventRecord(a, s), "event");
    for (u32 r = 0; r < reps; ++r) f();
    hip_check(hipEventRecord(b, s), "event");
    hip_check(hipEventSynchronize(b), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, b), "event");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

void Context::set_tuning(const std::string& key, int64_t value) {
    if (key == "fail_alloc_after") {  // a test hook (negative = off); captured plans stay valid
        fail_alloc_after = value;
        return;
    }
    if (key == "fail_launch_after") {  // test hook, as fail_alloc_after
        fail_launch_after = value;
        return;
    }
    if (key == "fail_upload") {  // test hook: 1 = every upload fails
        fail_upload = value != 0;
        return;
    }
    if (value < 0) fail(CKH_ERR_INVALID_PARAMS, "tuning values are non-negative");
    if (key == "ntt_chunk_fwd")
        ntt_chunk_fwd = (u32)value;
    else if (key == "ntt_chunk_inv")
        ntt_chunk_inv = (u32)value;
    else if (key == "graphs")
        use_graphs = value != 0;
    else if (key == "conv_variant")
        conv_variant = (u32)std::min<int64_t>(value, 2);
    else if (key == "conv_variant_md")
        conv_variant_md = (u32)std::min<int64_t>(value, 3);
    else if (key == "fuse_ip")
        fuse_ip = value != 0;
    else if (key == "fuse_fin")
        fuse_fin = value != 0;
    else if (key == "conv_yform")
        conv_yform = (u32)std::min<int64_t>(value, 7);
    else if (key == "small_limbs")
        small_limbs = (u32)std::min<int64_t>(value, 1u << 30);
    else if (key == "conv_variant_small")
        conv_variant_s = (u32)std::min<int64_t>(value, 2);
    else if (key == "conv_tg_small")
        conv_tg_s = (u32)std::min<int64_t>(std::max<int64_t>(value, 1), 64);
    else if (key == "fuse_ip_small")
        fuse_ip_s = value != 0;
    else if (key == "conv_tg")
        conv_tg = (u32)std::min<int64_t>(std::max<int64_t>(value, 1), 64);
    else if (key == "graph_cache")
        graph_cache_max = (u32)std::max<int64_t>(value, 1);
    else if (key == "pool") {
        // 1: the context's block cache; 0: the device's default stream-ordered pool (its frees are
        // ordered after every stream that used the buffer, as the cache's are).
        if (value > 1) fail(CKH_ERR_INVALID_PARAMS, "pool mode is 0 or 1");
        synchronize_all();
        pool_mode = (u32)value;
    }
    else if (key == "ntt_fused")
        ntt_fused = value != 0;
    else if (key == "fused_lag")
        fused.lag = (u32)std::min<int64_t>(std::max<int64_t>(value, 1), 64);
    else if (key == "hoisted_batch")
        hoisted_batch = (u32)std::min<int64_t>(std::max<int64_t>(value, 1), 64);
    else if (key == "fused_wg")  // workgroups per CU of the one-launch transforms
        fused.grid = cu_count * (u32)std::min<int64_t>(std::max<int64_t>(value, 1), 8);
    else if (key == "service_streams")
        service_streams = (u32)std::min<int64_t>(std::max<int64_t>(value, 1), SERVICE_STREAMS);
    else
        fail(CKH_ERR_INVALID_PARAMS, "unknown tuning key " + key);
    clear_plans();
}

int64_t Context::get_tuning(const std::string& key) const {
    if (key == "ntt_chunk_fwd") return ntt_chunk_fwd;
    if (key == "ntt_chunk_inv") return ntt_chunk_inv;
    if (key == "graphs") return use_graphs;
    if (key == "graph_cache") return graph_cache_max;
    if (key == "conv_variant") return conv_variant;
    if (key == "conv_variant_md") return conv_variant_md;
    if (key == "fuse_ip") return fuse_ip;
    if (key == "fuse_fin") return fuse_fin;
    if (key == "conv_yform") return conv_yform;
    if (key == "small_limbs") return small_limbs;
    if (key == "conv_variant_small") return conv_variant_s;
    if (key == "conv_tg_small") return conv_tg_s;
    if (key == "fuse_ip_small") return fuse_ip_s;
    if (key == "conv_tg") return conv_tg;
    if (key == "pool") return pool_mode;
    if (key == "service_streams") return service_streams;
    if (key == "ntt_fused") return fused_usable() ? 1 : 0;
    if (key == "fused_lag") return fused.lag;
    if (key == "hoisted_batch") return hoisted_batch;
    if (key == "stat_mallocs") return (int64_t)stat_mallocs.load();
    if (key == "stat_malloc_us") return (int64_t)stat_malloc_us.load();
    if (key == "stat_oom") return (int64_t)stat_oom.load();
    if (key == "stat_cached_bytes") {
        std::lock_guard<std::mutex> lk(const_cast<std::mutex&>(cache_mu_));
        return (int64_t)cache_bytes_;
    }
    if (key == "fused_wg") return cu_count ? fused.grid / cu_count : 0;
    if (key == "xcds") return fused.nx;
    if (key == "fail_alloc_after") return fail_alloc_after.load();
    if (key == "fail_launch_after") return fail_launch_after.load();
    if (key == "fail_upload") return fail_upload.load() ? 1 : 0;
    fail(CKH_ERR_INVALID_PARAMS, "unknown tuning key " + key);
}

// NTT chunk size: 16 polynomials of every Q and P prime (the limb mix of a key switch),
// transformed in place with each candidate; the fastest per direction is kept. Then the
// conversion kernel variant and the row-pass fusion, timed on a batched key switch.
std::string Context::autotune() {
    const u32 polys = 16, lp = nq + np;
    std::vector<PolyPtr> ps;
    for (u32 i = 0; i < polys; ++i) ps.push_back(poly_uniform(*this, nq - 1, np, false, 0x5eed + i, 0));
    std::vector<const Poly*> raw;
    for (auto& p : ps) raw.push_back(p.get());
    const u32 total = polys * lp;
    const u32 cands[] = {0, 16, 24, 32, 48, 64, 96, 128};
    std::ostringstream os;
    os << "{\"limbs\": " << total << ", \"candidates_ms\": {";
    const u32 keep_f = ntt_chunk_fwd, keep_i = ntt_chunk_inv;
    double best_f = 1e30, best_i = 1e30;
    u32 arg_f = keep_f, arg_i = keep_i;
    bool first = true;
    for (u32 c : cands) {
        if (c >= total) continue;
        ntt_chunk_fwd = ntt_chunk_inv = c;
        const double tf = time_ntt(*this, raw, true, 5), ti = time_ntt(*this, raw, false, 5);
        os << (first ? "" : ", ") << "\"" << c << "\": [" << tf << ", " << ti << "]";
        first = false;
        if (tf < best_f) best_f = tf, arg_f = c;
        if (ti < best_i) best_i = ti, arg_i = c;
    }
    ntt_chunk_fwd = arg_f;
    ntt_chunk_inv = arg_i;
    ps.clear();
    raw.clear();

    // Conversion kernel and row-pass fusion on a batched key switch of 12 polynomials at levels
    // spread over the chain under a synthetic key.
    const u32 top = nq - 1;
    KskPtr key = ksk_uniform(*this, top, 0, 0x5eed5eed);
    std::vector<PolyPtr> ds;
    std::vector<const Poly*> draw;
    for (u32 i = 0; i < 12; ++i) ds.push_back(poly_uniform(*this, top - (i * top) / 12, 0, true, 0x7eed + i, 0));
    for (auto& p : ds) draw.push_back(p.get());
    os << "}, \"ks_candidates_ms\": " << autotune_ks(draw, *key);
    os << ", \"small_candidates_ms\": " << autotune_small(*ds[0], *key);
    os << ", \"ntt_chunk_fwd\": " << ntt_chunk_fwd << ", \"ntt_chunk_inv\": " << ntt_chunk_inv << ", \"conv_variant\": " << conv_variant
       << ", \"conv_variant_md\": " << conv_variant_md << ", \"fuse_ip\": " << fuse_ip << ", \"small\": [" << conv_variant_s << ", " << conv_tg_s << ", " << fuse_ip_s
       << "]}";
    return os.str();
}

// The small-batch schedule (conversion variant, targets per workgroup, row-pass fusion), timed on
// one switch of d under key with captured launches, as a caller's single switch replays them.
std::string Context::autotune_small(const Poly& d, const Ksk& key) {
    const u32 keep_small = small_limbs;
    small_limbs = std::max(small_limbs, d.level + 1);  // d's switch takes the small schedule
    hipStream_t s = stream();
    auto time_one = [&]() {
        clear_plans();
        for (u32 r = 0; r < 3; ++r) (void)ks_switch(*this, d, key);  // direct run, capture, replay
        return time_on(s, 10, [&] { (void)ks_switch(*this, d, key); });
    };
    std::ostringstream os;
    os << "{";
    double best = 1e30;
    u32 bv = conv_variant_s, bt = conv_tg_s, bf = fuse_ip_s;
    bool first = true;
    for (u32 var : {0u, 1u, 2u})
        for (u32 tg : {4u, 8u, 16u, 24u})
            for (u32 fuse : {0u, 1u}) {
                if (fuse && !fuse_ip_possible()) continue;
                if (var == 2 && (tg != 8 || !conv.mf_ok)) continue;
                conv_variant_s = var;
                conv_tg_s = tg;
                fuse_ip_s = fuse;
                const double t = time_one();
                os << (first ? "" : ", ") << "\"conv" << var << "_tg" << tg << (fuse ? "_fused" : "") << "\": " << t;
                first = false;
                if (t < best) best = t, bv = var, bt = tg, bf = fuse;
            }
    conv_variant_s = bv;
    conv_tg_s = bt;
    fuse_ip_s = bf;
    small_limbs = keep_small;
    clear_plans();
    os << "}";
    return os.str();
}

// Times every (conversion variant, targets per workgroup, row-pass fusion) choice on the batched key switch of `ds`
// under `key` (direct launches, one warm-up and three timed runs each), keeps the fastest, then
// the inverse-NTT chunk against none, and returns the timings as a JSON object.
std::string Context::autotune_ks(const std::vector<const Poly*>& ds, const Ksk& key) {
    if (ds.empty()) fail(CKH_ERR_INVALID_PARAMS, "nothing to tune on");
    const u32 keep_graphs = use_graphs;
    use_graphs = 0;
    hipStream_t s = stream();
    auto time_ks = [&]() {
        clear_plans();
        (void)ks_switch_batch(*this, ds, key);  // warm
        return time_on(s, 3, [&] { (void)ks_switch_batch(*this, ds, key); });
    };
    std::ostringstream os;
    os << "{";
    double best = 1e30;
    u32 best_var = conv_variant, best_fuse = fuse_ip, best_tg = conv_tg;
    bool first = true;
    for (u32 var : {0u, 1u, 2u})
        for (u32 tg : {8u, 16u, 24u})
            for (u32 fuse : {0u, 1u}) {
                if (fuse && !fuse_ip_possible()) continue;
                if (var == 2 && (tg != 8 || !conv.mf_ok)) continue;  // one matrix-core schedule
                conv_variant = var;
                conv_tg = tg;
                fuse_ip = fuse;
                const double t = time_ks();
                os << (first ? "" : ", ") << "\"conv" << var << "_tg" << tg << (fuse ? "_fused" : "") << "\": " << t;
                first = false;
                if (t < best) best = t, best_var = var, best_fuse = fuse, best_tg = tg;
            }
    conv_variant = best_var;
    conv_tg = best_tg;
    fuse_ip = best_fuse;
    // mod_down's conversions (60-bit sources) on a variant of their own, against following
    // conv_variant (3); `best` was timed with the current choice.
    u32 best_md = conv_variant_md;
    for (u32 md : {3u, 0u, 1u, 2u}) {
        if (md == best_md || (md == 2 && !conv.mf_ok) || md == conv_variant) continue;
        conv_variant_md = md;
        const double t = time_ks();
        os << ", \"md" << md << "\": " << t;
        if (t < best) best = t, best_md = md;
    }
    conv_variant_md = best_md;
    // The inverse-NTT chunk picked on plain transforms, against no chunking, on the switch
    // itself (its mod_down INTT runs over P limbs grouped by prime, which changes what pays).
    u32 best_chunk = ntt_chunk_inv;
    if (ntt_chunk_inv != 0) {
        for (u32 chunk : {ntt_chunk_inv, 0u}) {
            ntt_chunk_inv = chunk;
            const double t = chunk == best_chunk ? best : time_ks();
            if (chunk != best_chunk) os << ", \"chunk_inv" << chunk << "\": " << t;
            if (t < best) best = t, best_chunk = chunk;
        }
    }
    ntt_chunk_inv = best_chunk;
    use_graphs = keep_graphs;
    clear_plans();
    os << "}";
    return os.str();
}

}  // namespace ckh
