// The project's autotuner: picks launch schedules per context by timing candidates on the
// device (HIP events on the context's stream) with synthetic limbs of the context's primes.
#include <algorithm>
#include <sstream>

#include "evaluator.h"

namespace ckh {

void Context::set_tuning(const std::string& key, int64_t value) {
    if (value < 0) fail(CKH_ERR_INVALID_PARAMS, "tuning values are non-negative");
    if (key == "ntt_chunk_fwd")
        ntt_chunk_fwd = (u32)value;
    else if (key == "ntt_chunk_inv")
        ntt_chunk_inv = (u32)value;
    else if (key == "graphs")
        use_graphs = value != 0;
    else if (key == "graph_cache")
        graph_cache_max = (u32)std::max<int64_t>(value, 1);
    else
        fail(CKH_ERR_INVALID_PARAMS, "unknown tuning key " + key);
    clear_plans();
}

// NTT chunk size: 16 polynomials of every Q and P prime (the limb mix of a key switch),
// transformed in place with each candidate; the fastest per direction is kept.
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
    clear_plans();
    os << "}, \"ntt_chunk_fwd\": " << arg_f << ", \"ntt_chunk_inv\": " << arg_i << "}";
    return os.str();
}

}  // namespace ckh
