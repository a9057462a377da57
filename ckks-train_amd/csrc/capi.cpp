// The C ABI (include/ckh.h) over the C++ runtime: handle types, status mapping and the
// calling thread's last error message.
#include <atomic>
#include <cstring>
#include <new>
#include <string>

#include "evaluator.h"
#include "hostmath.h"
#include "server.h"

using namespace ckh;

struct ckh_context {
    std::shared_ptr<Context> c;
};
struct ckh_poly {
    std::atomic<int> rc{1};
    PolyPtr p;
};
struct ckh_ksk {
    std::atomic<int> rc{1};
    KskPtr k;
};
struct ckh_ct {
    std::atomic<int> rc{1};
    CtPtr c;
};
struct ckh_evaluator {
    std::shared_ptr<Context> keep;
    // Shared with the serving loop's requests: an evaluator destroyed by its handle while one of
    // its requests is pending or in flight lives (with its keys) until that request completes.
    std::shared_ptr<Evaluator> e;
};
struct ckh_server {
    std::unique_ptr<Server> s;
};

namespace {
thread_local std::string tl_err;

template <class F>
ckh_status guarded(F&& f) {
    try {
        f();
        return CKH_OK;
    } catch (const Error& e) {
        tl_err = e.msg;
        return e.status;
    } catch (const std::bad_alloc&) {
        tl_err = "host out of memory";
        return CKH_ERR_DEVICE;
    } catch (const std::exception& e) {
        tl_err = e.what();
        return CKH_ERR_DEVICE;
    } catch (...) {
        tl_err = "unknown internal error";
        return CKH_ERR_DEVICE;
    }
}

Context& ctx_of(ckh_context* c) {
    if (!c || !c->c) fail(CKH_ERR_INVALID_PARAMS, "null context");
    hip_check(hipSetDevice(c->c->device), "hipSetDevice");
    return *c->c;
}

Evaluator& ev_of(ckh_evaluator* e) {
    if (!e || !e->e) fail(CKH_ERR_INVALID_PARAMS, "null evaluator");
    hip_check(hipSetDevice(e->e->ctx.device), "hipSetDevice");
    return *e->e;
}

const Poly& poly_of(const ckh_poly* p) {
    if (!p || !p->p) fail(CKH_ERR_INVALID_PARAMS, "null polynomial");
    return *p->p;
}

const Ct& ct_of(const ckh_ct* c) {
    if (!c || !c->c) fail(CKH_ERR_INVALID_PARAMS, "null ciphertext");
    return *c->c;
}

const Ksk& key_of(const ckh_ksk* k) {
    if (!k || !k->k) fail(CKH_ERR_INVALID_PARAMS, "null key");
    return *k->k;
}

ckh_poly* wrap(PolyPtr p) {
    auto* h = new ckh_poly();
    h->p = std::move(p);
    return h;
}

ckh_ct* wrap(CtPtr c) {
    auto* h = new ckh_ct();
    h->c = std::move(c);
    return h;
}

void need(const void* p, const char* what) {
    if (!p) fail(CKH_ERR_INVALID_PARAMS, std::string("null ") + what);
}
}  // namespace

extern "C" {

const char* ckh_last_error(void) { return tl_err.c_str(); }

ckh_status ckh_params_primes(uint32_t log_n, const uint32_t* q_bits, uint32_t nq, const uint32_t* p_bits,
                             uint32_t np, uint64_t* q_out, uint64_t* p_out) {
    return guarded([&] {
        if (log_n < 1 || log_n > 17) fail(CKH_ERR_INVALID_PARAMS, "log_n outside [1, 17]");
        if (nq == 0 || np == 0) fail(CKH_ERR_INVALID_PARAMS, "q_bits and p_bits must be non-empty");
        need(q_bits, "q_bits");
        need(p_bits, "p_bits");
        need(q_out, "q_out");
        need(p_out, "p_out");
        std::vector<uint32_t> all(q_bits, q_bits + nq);
        all.insert(all.end(), p_bits, p_bits + np);
        for (uint32_t b : all)
            if (b < 2 || b > 60) fail(CKH_ERR_INVALID_PARAMS, "prime bit size outside 2..=60");
        std::vector<uint64_t> primes(all.size());
        for (uint32_t b = 2; b <= 60; ++b) {
            uint64_t count = 0;
            for (uint32_t x : all) count += x == b;
            if (!count) continue;
            std::vector<uint64_t> pool;
            if (!hm::ntt_primes(b, 1ull << log_n, count, pool)) fail(CKH_ERR_INVALID_PARAMS, "not enough primes");
            uint64_t next = 0;
            for (size_t i = 0; i < all.size(); ++i)
                if (all[i] == b) primes[i] = pool[count - 1 - next++];
        }
        std::memcpy(q_out, primes.data(), nq * 8);
        std::memcpy(p_out, primes.data() + nq, np * 8);
    });
}

uint64_t ckh_galois_elt(uint32_t log_n, int32_t step) { return galois_elt(1ull << log_n, step); }

uint64_t ckh_conjugation_elt(uint32_t log_n) { return (2ull << log_n) - 1; }

ckh_status ckh_device_open(int ordinal, const ckh_params* params, uint32_t n_streams, ckh_context** out) {
    return guarded([&] {
        need(params, "params");
        need(out, "out");
        auto h = new ckh_context();
        try {
            h->c = std::make_shared<Context>(ordinal, *params, n_streams);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

void ckh_device_close(ckh_context* ctx) { delete ctx; }

ckh_status ckh_set_stream(ckh_context* ctx, uint32_t index) {
    return guarded([&] { ctx_of(ctx).set_stream(index); });
}

void* ckh_stream_handle(ckh_context* ctx) { return ctx && ctx->c ? (void*)ctx->c->stream() : nullptr; }

ckh_status ckh_autotune(ckh_context* ctx, char* report, size_t report_len) {
    return guarded([&] {
        std::string r = ctx_of(ctx).autotune();
        if (report && report_len) {
            std::strncpy(report, r.c_str(), report_len - 1);
            report[report_len - 1] = 0;
        }
    });
}

ckh_status ckh_autotune_ks(ckh_context* ctx, uint32_t n, const ckh_poly* const* d, const ckh_ksk* key, char* report,
                           size_t report_len) {
    return guarded([&] {
        need(d, "d");
        std::vector<const Poly*> ds;
        for (uint32_t i = 0; i < n; ++i) ds.push_back(&poly_of(d[i]));
        std::string r = ctx_of(ctx).autotune_ks(ds, key_of(key));
        if (report && report_len) {
            std::strncpy(report, r.c_str(), report_len - 1);
            report[report_len - 1] = 0;
        }
    });
}

ckh_status ckh_set_tuning(ckh_context* ctx, const char* key, int64_t value) {
    return guarded([&] {
        need(key, "key");
        ctx_of(ctx).set_tuning(key, value);
    });
}

ckh_status ckh_get_tuning(ckh_context* ctx, const char* key, int64_t* value) {
    return guarded([&] {
        need(key, "key");
        need(value, "value");
        *value = ctx_of(ctx).get_tuning(key);
    });
}

ckh_status ckh_graph_stats(ckh_context* ctx, uint64_t* captures, uint64_t* launches) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        if (captures) *captures = c.graph_captures;
        if (launches) *launches = c.graph_launches;
    });
}

ckh_status ckh_pool_used_memory(ckh_context* ctx, uint64_t* bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        *bytes = ctx_of(ctx).pool_used_memory();
    });
}

ckh_status ckh_set_pool_release_threshold(ckh_context* ctx, uint64_t bytes) {
    return guarded([&] { ctx_of(ctx).set_pool_release_threshold(bytes); });
}

ckh_status ckh_trim_pool(ckh_context* ctx, uint64_t keep_bytes) {
    return guarded([&] { ctx_of(ctx).trim_pool(keep_bytes); });
}

ckh_status ckh_synchronize(ckh_context* ctx) {
    return guarded([&] {
        ctx_of(ctx).synchronize_all();
    });
}

ckh_status ckh_memory_info(ckh_context* ctx, uint64_t* live_bytes, uint64_t* free_bytes) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        size_t fr = 0, tot = 0;
        hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        if (live_bytes) *live_bytes = c.live_bytes.load();
        if (free_bytes) *free_bytes = fr;
    });
}

ckh_status ckh_poly_upload(ckh_context* ctx, const uint64_t* limbs, uint32_t level, uint32_t p_limbs, int ntt_form,
                           ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_upload(ctx_of(ctx), limbs, level, p_limbs, ntt_form != 0));
    });
}

ckh_status ckh_poly_download(ckh_context* ctx, const ckh_poly* poly, uint64_t* out) {
    return guarded([&] { poly_download(ctx_of(ctx), poly_of(poly), out); });
}

ckh_status ckh_poly_digest(ckh_context* ctx, const ckh_poly* poly, uint64_t* digest) {
    return guarded([&] {
        need(digest, "digest");
        *digest = poly_digest(ctx_of(ctx), poly_of(poly));
    });
}

ckh_status ckh_poly_info(const ckh_poly* poly, uint32_t* level, uint32_t* p_limbs, int* ntt_form) {
    return guarded([&] {
        const Poly& p = poly_of(poly);
        if (level) *level = p.level;
        if (p_limbs) *p_limbs = p.p_limbs;
        if (ntt_form) *ntt_form = p.ntt;
    });
}

ckh_poly* ckh_poly_retain(ckh_poly* poly) {
    if (poly) poly->rc++;
    return poly;
}

void ckh_poly_release(ckh_poly* poly) {
    if (poly && --poly->rc == 0) {
        if (poly->p && poly->p->buf) (void)hipSetDevice(poly->p->buf->ctx->device);
        delete poly;
    }
}

ckh_status ckh_poly_uniform(ckh_context* ctx, uint32_t level, uint32_t p_limbs, int ntt_form, uint64_t seed,
                            uint64_t stream0, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_uniform(ctx_of(ctx), level, p_limbs, ntt_form != 0, seed, stream0));
    });
}

ckh_status ckh_ntt(ckh_context* ctx, const ckh_poly* in, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_transform(ctx_of(ctx), poly_of(in), true));
    });
}

ckh_status ckh_intt(ckh_context* ctx, const ckh_poly* in, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_transform(ctx_of(ctx), poly_of(in), false));
    });
}

ckh_status ckh_automorphism(ckh_context* ctx, const ckh_poly* in, uint64_t galois_elt, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_automorphism(ctx_of(ctx), poly_of(in), galois_elt));
    });
}

ckh_status ckh_time_ntt(ckh_context* ctx, ckh_poly* const* polys, uint32_t n, int forward, uint32_t reps,
                        double* ms_per_launch) {
    return guarded([&] {
        need(polys, "polys");
        need(ms_per_launch, "ms_per_launch");
        std::vector<const Poly*> v;
        for (uint32_t i = 0; i < n; ++i) v.push_back(&poly_of(polys[i]));
        *ms_per_launch = time_ntt(ctx_of(ctx), v, forward != 0, reps);
    });
}

ckh_status ckh_time_copy(ckh_context* ctx, uint64_t bytes, uint32_t reps, double* ms_per_launch) {
    return guarded([&] {
        need(ms_per_launch, "ms_per_launch");
        *ms_per_launch = time_copy(ctx_of(ctx), bytes, reps);
    });
}

ckh_status ckh_rescale_poly(ckh_context* ctx, const ckh_poly* in, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_rescale(ctx_of(ctx), poly_of(in)));
    });
}

ckh_status ckh_raise_from_level0(ckh_context* ctx, const ckh_poly* in, uint32_t top, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_raise(ctx_of(ctx), poly_of(in), top));
    });
}

ckh_status ckh_mod_down(ckh_context* ctx, const ckh_poly* qp, ckh_poly** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(poly_mod_down(ctx_of(ctx), poly_of(qp)));
    });
}

ckh_status ckh_ksk_upload(ckh_context* ctx, const uint64_t* limbs, uint32_t max_level, uint64_t galois_elt,
                          ckh_ksk** out) {
    return guarded([&] {
        need(out, "out");
        need(limbs, "limbs");
        auto* h = new ckh_ksk();
        h->k = ksk_upload(ctx_of(ctx), limbs, max_level, galois_elt);
        *out = h;
    });
}

ckh_status ckh_ksk_uniform(ckh_context* ctx, uint32_t max_level, uint64_t galois_elt, uint64_t seed, ckh_ksk** out) {
    return guarded([&] {
        need(out, "out");
        auto* h = new ckh_ksk();
        h->k = ksk_uniform(ctx_of(ctx), max_level, galois_elt, seed);
        *out = h;
    });
}

ckh_status ckh_ksk_info(const ckh_ksk* key, uint32_t* max_level, uint64_t* galois_elt, uint64_t* bytes) {
    return guarded([&] {
        const Ksk& k = key_of(key);
        if (max_level) *max_level = k.max_level;
        if (galois_elt) *galois_elt = k.galois;
        if (bytes) *bytes = (uint64_t)k.parts * 2 * k.poly_words() * 8;
    });
}

void ckh_ksk_release(ckh_ksk* key) {
    if (key && --key->rc == 0) delete key;
}

ckh_status ckh_decompose(ckh_context* ctx, const ckh_poly* d, ckh_poly** digits, uint32_t* n_digits) {
    return guarded([&] {
        need(digits, "digits");
        auto v = ks_decompose(ctx_of(ctx), poly_of(d));
        for (size_t j = 0; j < v.size(); ++j) digits[j] = wrap(v[j]);
        if (n_digits) *n_digits = (uint32_t)v.size();
    });
}

ckh_status ckh_apply_key(ckh_context* ctx, ckh_poly* const* digits, uint32_t n_digits, const ckh_ksk* key,
                         ckh_poly** kb, ckh_poly** ka) {
    return guarded([&] {
        need(digits, "digits");
        need(kb, "kb");
        need(ka, "ka");
        std::vector<PolyPtr> v;
        for (uint32_t j = 0; j < n_digits; ++j) v.push_back(std::make_shared<Poly>(poly_of(digits[j])));
        auto r = ks_apply(ctx_of(ctx), v, key_of(key));
        *kb = wrap(r.first);
        *ka = wrap(r.second);
    });
}

ckh_status ckh_key_switch(ckh_context* ctx, const ckh_poly* d, const ckh_ksk* key, ckh_poly** kb, ckh_poly** ka) {
    return guarded([&] {
        need(kb, "kb");
        need(ka, "ka");
        auto r = ks_switch(ctx_of(ctx), poly_of(d), key_of(key));
        *kb = wrap(r.first);
        *ka = wrap(r.second);
    });
}

ckh_status ckh_key_switch_batch(ckh_context* ctx, uint32_t n, const ckh_poly* const* d, const ckh_ksk* key,
                                ckh_poly** kb, ckh_poly** ka) {
    return guarded([&] {
        need(d, "d");
        need(kb, "kb");
        need(ka, "ka");
        std::vector<const Poly*> v;
        for (uint32_t i = 0; i < n; ++i) v.push_back(&poly_of(d[i]));
        auto r = ks_switch_batch(ctx_of(ctx), v, key_of(key));
        for (uint32_t i = 0; i < n; ++i) {
            kb[i] = wrap(r[i].first);
            ka[i] = wrap(r[i].second);
        }
    });
}

ckh_status ckh_key_switch_batch_static(ckh_context* ctx, uint32_t n, const ckh_poly* const* d, const ckh_ksk* key,
                                       ckh_poly* const* kb, ckh_poly* const* ka) {
    return guarded([&] {
        need(d, "d");
        need(kb, "kb");
        need(ka, "ka");
        std::vector<const Poly*> v;
        std::vector<Poly*> ob, oa;
        for (uint32_t i = 0; i < n; ++i) {
            v.push_back(&poly_of(d[i]));
            ob.push_back(const_cast<Poly*>(&poly_of(kb[i])));
            oa.push_back(const_cast<Poly*>(&poly_of(ka[i])));
        }
        ks_switch_batch_static(ctx_of(ctx), v, key_of(key), ob, oa);
    });
}

ckh_status ckh_ct_upload(ckh_context* ctx, const uint64_t* polys, uint32_t n_comp, uint32_t level, double scale,
                         ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        need(polys, "polys");
        if (n_comp < 1) fail(CKH_ERR_INVALID_PARAMS, "a ciphertext needs at least one component");
        Context& c = ctx_of(ctx);
        auto ct = std::make_shared<Ct>();
        for (uint32_t i = 0; i < n_comp; ++i)
            ct->comps.push_back(poly_upload(c, polys + (uint64_t)i * (level + 1) * c.n, level, 0, true));
        ct->scale = scale;
        *out = wrap(ct);
    });
}

ckh_status ckh_ct_from_polys(ckh_context* ctx, ckh_poly* const* polys, uint32_t n_comp, double scale,
                             ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        need(polys, "polys");
        Context& c = ctx_of(ctx);
        auto ct = std::make_shared<Ct>();
        for (uint32_t i = 0; i < n_comp; ++i) {
            const Poly& p = poly_of(polys[i]);
            if (p.buf->ctx != &c || p.p_limbs || !p.ntt)
                fail(CKH_ERR_INVALID_PARAMS, "components must be NTT-form Q polynomials of this context");
            if (i && p.level != ct->comps[0]->level) fail(CKH_ERR_LEVEL_MISMATCH, "components at different levels");
            ct->comps.push_back(polys[i]->p);
        }
        if (ct->comps.empty()) fail(CKH_ERR_INVALID_PARAMS, "a ciphertext needs at least one component");
        ct->scale = scale;
        *out = wrap(ct);
    });
}

ckh_status ckh_ct_download(ckh_context* ctx, const ckh_ct* ct, uint64_t* out) {
    return guarded([&] {
        Context& c = ctx_of(ctx);
        const Ct& x = ct_of(ct);
        for (size_t i = 0; i < x.comps.size(); ++i)
            poly_download(c, *x.comps[i], out + i * (uint64_t)(x.level() + 1) * c.n);
    });
}

ckh_status ckh_ct_info(const ckh_ct* ct, uint32_t* n_comp, uint32_t* level, double* scale) {
    return guarded([&] {
        const Ct& x = ct_of(ct);
        if (n_comp) *n_comp = (uint32_t)x.comps.size();
        if (level) *level = x.level();
        if (scale) *scale = x.scale;
    });
}

ckh_poly* ckh_ct_component(const ckh_ct* ct, uint32_t index) {
    if (!ct || !ct->c || index >= ct->c->comps.size()) return nullptr;
    return wrap(ct->c->comps[index]);
}

ckh_ct* ckh_ct_retain(ckh_ct* ct) {
    if (ct) ct->rc++;
    return ct;
}

void ckh_ct_release(ckh_ct* ct) {
    if (ct && --ct->rc == 0) delete ct;
}

ckh_status ckh_evaluator_create(ckh_context* ctx, const ckh_ksk* relin, const ckh_ksk* const* galois,
                                uint32_t n_galois, ckh_evaluator** out) {
    return guarded([&] {
        need(out, "out");
        Context& c = ctx_of(ctx);
        std::vector<KskPtr> g;
        for (uint32_t i = 0; i < n_galois; ++i) {
            const Ksk& k = key_of(galois[i]);
            if (k.galois == 0) fail(CKH_ERR_INVALID_PARAMS, "galois key list holds a relinearization key");
            if (k.ctx != &c) fail(CKH_ERR_INVALID_PARAMS, "key is not over this context");
            g.push_back(galois[i]->k);
        }
        KskPtr r;
        if (relin) {
            const Ksk& k = key_of(relin);
            if (k.galois != 0) fail(CKH_ERR_INVALID_PARAMS, "relinearization key has a Galois element");
            if (k.ctx != &c) fail(CKH_ERR_INVALID_PARAMS, "key is not over this context");
            r = relin->k;
        }
        auto* h = new ckh_evaluator();
        h->keep = ctx->c;
        h->e = std::make_shared<Evaluator>(c, r, g);
        *out = h;
    });
}

ckh_status ckh_evaluator_fork(ckh_evaluator* ev, ckh_evaluator** out) {
    return guarded([&] {
        need(out, "out");
        Evaluator& e = ev_of(ev);
        std::vector<KskPtr> g;
        for (auto& kv : e.galois) g.push_back(kv.second);
        auto* h = new ckh_evaluator();
        h->keep = ev->keep;
        h->e = std::make_shared<Evaluator>(e.ctx, e.relin, g);
        h->e->config = e.config;
        *out = h;
    });
}

ckh_status ckh_evaluator_set_config(ckh_evaluator* ev, uint32_t hoist, uint32_t bsgs, uint32_t lazy_relin) {
    return guarded([&] {
        EvalConfig& c = ev_of(ev).config;
        c.hoist = hoist != 0;
        c.bsgs = bsgs != 0;
        c.lazy_relin = lazy_relin != 0;
    });
}

ckh_status ckh_evaluator_get_config(ckh_evaluator* ev, uint32_t* hoist, uint32_t* bsgs, uint32_t* lazy_relin) {
    return guarded([&] {
        need(hoist, "hoist");
        need(bsgs, "bsgs");
        need(lazy_relin, "lazy_relin");
        const EvalConfig& c = ev_of(ev).config;
        *hoist = c.hoist;
        *bsgs = c.bsgs;
        *lazy_relin = c.lazy_relin;
    });
}

void ckh_evaluator_destroy(ckh_evaluator* ev) { delete ev; }

ckh_status ckh_counter(ckh_evaluator* ev, ckh_op_counter* out) {
    return guarded([&] {
        need(out, "out");
        *out = ev_of(ev).counter();
    });
}

ckh_status ckh_reset_counter(ckh_evaluator* ev) {
    return guarded([&] { ev_of(ev).reset_counter(); });
}

ckh_status ckh_add_counts(ckh_evaluator* ev, const ckh_op_counter* counts) {
    return guarded([&] {
        need(counts, "counts");
        ev_of(ev).add_counts(*counts);
    });
}

int ckh_has_relin_key(ckh_evaluator* ev) { return ev && ev->e && ev->e->relin ? 1 : 0; }

int ckh_has_galois_key(ckh_evaluator* ev, int32_t step) { return ev && ev->e && ev->e->has_galois_key(step) ? 1 : 0; }

int ckh_owns(ckh_evaluator* ev, const ckh_ct* ct) { return ev && ev->e && ct && ct->c && ev->e->owns(*ct->c) ? 1 : 0; }

#define CT_OP(expr)          \
    return guarded([&] {     \
        need(out, "out");    \
        *out = wrap(expr);   \
    })

ckh_status ckh_add(ckh_evaluator* ev, const ckh_ct* a, const ckh_ct* b, ckh_ct** out) {
    CT_OP(ev_of(ev).add(ct_of(a), ct_of(b), false));
}
ckh_status ckh_sub(ckh_evaluator* ev, const ckh_ct* a, const ckh_ct* b, ckh_ct** out) {
    CT_OP(ev_of(ev).add(ct_of(a), ct_of(b), true));
}
ckh_status ckh_negate(ckh_evaluator* ev, const ckh_ct* a, ckh_ct** out) { CT_OP(ev_of(ev).negate(ct_of(a))); }
ckh_status ckh_add_plain(ckh_evaluator* ev, const ckh_ct* a, const ckh_poly* pt, double pt_scale, ckh_ct** out) {
    CT_OP(ev_of(ev).add_plain(ct_of(a), poly_of(pt), pt_scale));
}
ckh_status ckh_mul_plain(ckh_evaluator* ev, const ckh_ct* a, const ckh_poly* pt, double pt_scale, ckh_ct** out) {
    CT_OP(ev_of(ev).mul_plain(ct_of(a), poly_of(pt), pt_scale));
}
ckh_status ckh_mul_scalar(ckh_evaluator* ev, const ckh_ct* ct, double value, double target_scale, ckh_ct** out) {
    CT_OP(ev_of(ev).mul_scalar(ct_of(ct), value, target_scale));
}
ckh_status ckh_mul_plain_to_scale(ckh_evaluator* ev, const ckh_ct* a, const ckh_poly* pt, double pt_scale,
                                  double target_scale, ckh_ct** out) {
    CT_OP(ev_of(ev).mul_plain_to_scale(ct_of(a), poly_of(pt), pt_scale, target_scale));
}
ckh_status ckh_exact_product_scale(ckh_evaluator* ev, uint32_t level, double target_scale, double* out) {
    return guarded([&] {
        need(out, "out");
        *out = ev_of(ev).exact_product_scale(level, target_scale);
    });
}
ckh_status ckh_target_plain_scale(ckh_evaluator* ev, const ckh_ct* ct, double target_scale, double* out) {
    return guarded([&] {
        need(out, "out");
        *out = ev_of(ev).target_plain_scale(ct_of(ct), target_scale);
    });
}
ckh_status ckh_mod_raise(ckh_evaluator* ev, const ckh_ct* ct, const ckh_ksk* to_sparse, const ckh_ksk* to_dense,
                         ckh_ct** out) {
    CT_OP(ev_of(ev).mod_raise(ct_of(ct), key_of(to_sparse), key_of(to_dense)));
}
ckh_status ckh_mod_raise_main_key(ckh_evaluator* ev, const ckh_ct* ct, ckh_ct** out) {
    CT_OP(ev_of(ev).mod_raise_main_key(ct_of(ct)));
}
ckh_status ckh_mul_lazy(ckh_evaluator* ev, const ckh_ct* a, const ckh_ct* b, ckh_ct** out) {
    CT_OP(ev_of(ev).mul_lazy(ct_of(a), ct_of(b)));
}
ckh_status ckh_relin_lazy(ckh_evaluator* ev, const ckh_ct* sum, ckh_ct** out) {
    CT_OP(ev_of(ev).relin_lazy(ct_of(sum)));
}
ckh_status ckh_mul_no_relin(ckh_evaluator* ev, const ckh_ct* a, const ckh_ct* b, ckh_ct** out) {
    CT_OP(ev_of(ev).mul_no_relin(ct_of(a), ct_of(b)));
}
ckh_status ckh_relinearize(ckh_evaluator* ev, const ckh_ct* ct, ckh_ct** out) {
    CT_OP(ev_of(ev).relinearize(ct_of(ct)));
}
ckh_status ckh_mul(ckh_evaluator* ev, const ckh_ct* a, const ckh_ct* b, ckh_ct** out) {
    CT_OP(ev_of(ev).mul(ct_of(a), ct_of(b)));
}
ckh_status ckh_rescale(ckh_evaluator* ev, const ckh_ct* ct, ckh_ct** out) { CT_OP(ev_of(ev).rescale(ct_of(ct))); }
ckh_status ckh_mod_switch_to_level(ckh_evaluator* ev, const ckh_ct* ct, uint32_t level, ckh_ct** out) {
    CT_OP(ev_of(ev).mod_switch_to_level(ct_of(ct), level));
}
ckh_status ckh_relabel(ckh_evaluator* ev, const ckh_ct* ct, double scale, ckh_ct** out) {
    CT_OP(ev_of(ev).relabel(ct_of(ct), scale));
}
ckh_status ckh_rotate(ckh_evaluator* ev, const ckh_ct* ct, int32_t steps, ckh_ct** out) {
    CT_OP(ev_of(ev).rotate(ct_of(ct), steps));
}
ckh_status ckh_conjugate(ckh_evaluator* ev, const ckh_ct* ct, ckh_ct** out) {
    CT_OP(ev_of(ev).conjugate(ct_of(ct)));
}

ckh_status ckh_hoisted_rotate(ckh_evaluator* ev, const ckh_ct* ct, const int32_t* steps, uint32_t n_steps,
                              ckh_ct** outs) {
    return guarded([&] {
        need(outs, "outs");
        if (n_steps) need(steps, "steps");
        auto r = ev_of(ev).hoisted_rotate(ct_of(ct), std::vector<int32_t>(steps, steps + n_steps));
        for (uint32_t i = 0; i < n_steps; ++i) outs[i] = wrap(r[i]);
    });
}

ckh_status ckh_matvec_plain(ckh_evaluator* ev, const ckh_ct* x, const ckh_poly* const* entries, double pt_scale,
                            uint32_t baby, uint32_t giant, uint32_t stride, int32_t offset, ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        need(entries, "entries");
        std::vector<const Poly*> v;
        for (uint64_t i = 0; i < (uint64_t)baby * giant; ++i) v.push_back(entries[i] ? &poly_of(entries[i]) : nullptr);
        *out = wrap(ev_of(ev).matvec_plain(ct_of(x), v, pt_scale, baby, giant, stride, offset));
    });
}

ckh_status ckh_matvec_enc(ckh_evaluator* ev, const ckh_ct* x, const ckh_ct* const* entries, uint32_t baby,
                          uint32_t giant, uint32_t stride, int32_t offset, ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        need(entries, "entries");
        std::vector<const Ct*> v;
        for (uint64_t i = 0; i < (uint64_t)baby * giant; ++i) v.push_back(entries[i] ? &ct_of(entries[i]) : nullptr);
        *out = wrap(ev_of(ev).matvec_enc(ct_of(x), v, baby, giant, stride, offset));
    });
}

ckh_status ckh_matvec_plain_sum(ckh_evaluator* ev, uint32_t n_terms, const ckh_ct* const* xs,
                                const ckh_poly* const* const* entries, double pt_scale, uint32_t baby,
                                uint32_t giant, uint32_t stride, int32_t offset, ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        if (n_terms) {
            need(xs, "xs");
            need(entries, "entries");
        }
        std::vector<const Ct*> x;
        std::vector<std::vector<const Poly*>> e(n_terms);
        for (uint32_t t = 0; t < n_terms; ++t) {
            x.push_back(&ct_of(xs[t]));
            need(entries[t], "entries");
            for (uint64_t j = 0; j < (uint64_t)baby * giant; ++j)
                e[t].push_back(entries[t][j] ? &poly_of(entries[t][j]) : nullptr);
        }
        *out = wrap(ev_of(ev).matvec_plain_sum(x, e, pt_scale, baby, giant, stride, offset));
    });
}

static std::vector<LinearStage> linear_stages(uint32_t n, const ckh_linear_stage* st) {
    if (n) need(st, "stages");
    std::vector<LinearStage> out;
    for (uint32_t k = 0; k < n; ++k) {
        const ckh_linear_stage& s = st[k];
        LinearStage lt;
        lt.slot_to_coeff = s.slot_to_coeff != 0;
        lt.level = s.level;
        lt.scale = s.scale;
        lt.shape = {s.shape.baby, s.shape.giant, s.shape.stride, s.shape.offset};
        lt.pt_scale = s.pt_scale;
        need(s.diags, "diags");
        for (uint64_t j = 0; j < (uint64_t)s.shape.baby * s.shape.giant; ++j)
            lt.diags.push_back(s.diags[j] ? &poly_of(s.diags[j]) : nullptr);
        lt.has_second = s.second != nullptr;
        lt.second_shape = {s.second_shape.baby, s.second_shape.giant, s.second_shape.stride, s.second_shape.offset};
        lt.second_pt_scale = s.second_pt_scale;
        if (lt.has_second)
            for (uint64_t j = 0; j < (uint64_t)s.second_shape.baby * s.second_shape.giant; ++j)
                lt.second.push_back(s.second[j] ? &poly_of(s.second[j]) : nullptr);
        out.push_back(std::move(lt));
    }
    return out;
}

ckh_status ckh_coeff_to_slot(ckh_evaluator* ev, const ckh_ct* ct, uint32_t n_stages, const ckh_linear_stage* stages,
                             ckh_ct** re, ckh_ct** im) {
    return guarded([&] {
        need(re, "re");
        need(im, "im");
        auto r = ev_of(ev).coeff_to_slot(ct_of(ct), linear_stages(n_stages, stages));
        *re = wrap(r.first);
        *im = wrap(r.second);
    });
}

ckh_status ckh_slot_to_coeff(ckh_evaluator* ev, const ckh_ct* re, const ckh_ct* im, uint32_t n_stages,
                             const ckh_linear_stage* stages, ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        *out = wrap(ev_of(ev).slot_to_coeff(ct_of(re), ct_of(im), linear_stages(n_stages, stages)));
    });
}

ckh_status ckh_matvec_enc_batch(ckh_evaluator* ev, uint32_t n, const ckh_ct* const* xs,
                                const ckh_ct* const* const* entries, const ckh_bsgs_shape* shapes, ckh_ct** outs) {
    return guarded([&] {
        need(outs, "outs");
        if (n) {
            need(xs, "xs");
            need(entries, "entries");
            need(shapes, "shapes");
        }
        std::vector<const Ct*> x;
        std::vector<std::vector<const Ct*>> e(n);
        std::vector<BsgsShape> sh;
        for (uint32_t i = 0; i < n; ++i) {
            x.push_back(&ct_of(xs[i]));
            sh.push_back({shapes[i].baby, shapes[i].giant, shapes[i].stride, shapes[i].offset});
            need(entries[i], "entries");
            for (uint64_t j = 0; j < (uint64_t)shapes[i].baby * shapes[i].giant; ++j)
                e[i].push_back(entries[i][j] ? &ct_of(entries[i][j]) : nullptr);
        }
        auto r = ev_of(ev).matvec_enc_batch(x, e, sh);
        for (uint32_t i = 0; i < n; ++i) outs[i] = wrap(r[i]);
    });
}

ckh_status ckh_matvec_enc_workspace_bytes(ckh_evaluator* ev, uint32_t level, uint32_t baby, uint32_t giant,
                                          uint64_t* bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        if (baby == 0 || giant == 0) fail(CKH_ERR_INVALID_PARAMS, "baby and giant must be positive");
        if (level >= ev_of(ev).ctx.nq) fail(CKH_ERR_INVALID_PARAMS, "level exceeds max_level");
        *bytes = ev_of(ev).matvec_enc_workspace_bytes(level, baby, giant);
    });
}

ckh_status ckh_relinearize_batch(ckh_evaluator* ev, uint32_t n, const ckh_ct* const* in, ckh_ct** out) {
    return guarded([&] {
        need(out, "out");
        if (n) need(in, "in");
        std::vector<const Ct*> v;
        for (uint32_t i = 0; i < n; ++i) v.push_back(&ct_of(in[i]));
        auto r = ev_of(ev).relinearize_batch(v);
        for (uint32_t i = 0; i < n; ++i) out[i] = wrap(r[i]);
    });
}

ckh_status ckh_server_create(const ckh_server_config* cfg, ckh_server** out) {
    return guarded([&] {
        need(cfg, "cfg");
        need(out, "out");
        auto* h = new ckh_server();
        h->s.reset(new Server(*cfg));
        *out = h;
    });
}

void ckh_server_destroy(ckh_server* srv) { delete srv; }

double ckh_server_now(ckh_server* srv) { return srv && srv->s ? srv->s->now() : 0.0; }

ckh_status ckh_server_submit(ckh_server* srv, const ckh_request* req, uint64_t* ticket) {
    return guarded([&] {
        need(srv, "server");
        need(req, "request");
        need(ticket, "ticket");
        auto op = std::make_shared<ServedOp>();
        op->kind = req->kind;
        ev_of(req->ev);  // a valid handle
        op->ev = req->ev->e;
        op->a = req->a ? req->a->c : nullptr;
        op->b = req->b ? req->b->c : nullptr;
        if (req->steps && req->n_steps) op->steps.assign(req->steps, req->steps + req->n_steps);
        op->t_arrival = req->arrival_s;
        op->inject_failure = req->inject_failure != 0;
        *ticket = srv->s->submit(op);
    });
}

ckh_status ckh_server_poll(ckh_server* srv, uint64_t ticket, int wait, int* finished, ckh_ct** outs, uint32_t max_outs,
                           uint32_t* n_outs, ckh_status* op_status, double* latency_ms) {
    return guarded([&] {
        need(srv, "server");
        need(finished, "finished");
        OpPtr op;
        *finished = srv->s->poll(ticket, wait != 0, op) ? 1 : 0;
        if (!*finished) return;
        if (op_status) *op_status = op->status;
        if (op->status != CKH_OK) tl_err = op->err;
        if (latency_ms) *latency_ms = (op->t_done - op->t_arrival) * 1e3;
        const uint32_t n = (uint32_t)std::min<size_t>(op->outs.size(), max_outs);
        if (n) need(outs, "outs");
        for (uint32_t i = 0; i < n; ++i) outs[i] = wrap(op->outs[i]);
        if (n_outs) *n_outs = (uint32_t)op->outs.size();
    });
}

ckh_status ckh_server_stats_get(ckh_server* srv, ckh_server_stats* out) {
    return guarded([&] {
        need(srv, "server");
        need(out, "out");
        *out = srv->s->stats();
    });
}

}  // extern "C"
