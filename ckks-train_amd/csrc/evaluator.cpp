// Polynomial operations and the HeOps evaluator over the device engine.
#include "evaluator.h"

#include <cmath>
#include <cstring>
#include <sstream>

#include "hostmath.h"

namespace ckh {

// ------------------------------------------------------------------ polynomials

// The calling thread's stream is about to queue work on this buffer: recorded so that the
// buffer's free is ordered after it (Buffer::note_use, Context::free_buffer).
static void use(const Context& c, const Poly& p) { p.buf->note_use(c.stream()); }
static void use(const Context& c, const Ksk& k) { k.buf->note_use(c.stream()); }

PolyPtr poly_alloc(Context& c, u32 level, u32 p_limbs, bool ntt) {
    if (level >= c.nq) fail(CKH_ERR_INVALID_PARAMS, "level above max_level");
    if (p_limbs != 0 && p_limbs != c.np) fail(CKH_ERR_INVALID_PARAMS, "p_limbs must be 0 or the special prime count");
    auto p = std::make_shared<Poly>();
    p->buf = c.alloc((u64)(level + 1 + p_limbs) * c.n);
    p->off = 0;
    p->level = level;
    p->p_limbs = p_limbs;
    p->ntt = ntt;
    return p;
}

static PolyPtr view(const BufPtr& b, u64 off, u32 level, u32 p_limbs, bool ntt) {
    auto p = std::make_shared<Poly>();
    p->buf = b;
    p->off = off;
    p->level = level;
    p->p_limbs = p_limbs;
    p->ntt = ntt;
    return p;
}

static u64 prime_of_limb(const Context& c, const Poly& p, u32 i) {
    return c.primes[i <= p.level ? i : c.nq + (i - p.level - 1)];
}

PolyPtr poly_upload(Context& c, const u64* host, u32 level, u32 p_limbs, bool ntt) {
    if (!host) fail(CKH_ERR_INVALID_PARAMS, "null limbs");
    c.tick_upload();
    PolyPtr p = poly_alloc(c, level, p_limbs, ntt);
    for (u32 i = 0; i < p->limbs(); ++i) {
        const u64 q = prime_of_limb(c, *p, i);
        for (u64 k = 0; k < c.n; ++k)
            if (host[i * c.n + k] >= q) fail(CKH_ERR_INVALID_PARAMS, "limb entry not reduced");
    }
    hipStream_t s = c.stream();
    c.h2d(p->data(), host, (u64)p->limbs() * c.n * 8, s);
    hip_check(hipStreamSynchronize(s), "upload");
    return p;
}

void poly_download(Context& c, const Poly& p, u64* host) {
    if (!host) fail(CKH_ERR_INVALID_PARAMS, "null output");
    use(c, p);
    hipStream_t s = c.stream();
    hip_check(hipMemcpyAsync(host, p.data(), (u64)p.limbs() * c.n * 8, hipMemcpyDeviceToHost, s), "download");
    hip_check(hipStreamSynchronize(s), "download");
    c.check_device_error(s);
}

u64 poly_digest(Context& c, const Poly& p) {
    use(c, p);
    hipStream_t s = c.stream();
    BufPtr acc = c.alloc(1, s);
    hip_check(hipMemsetAsync(acc->ptr, 0, 8, s), "memset");
    digest_launch(p.data(), (u64)p.limbs() * c.n, acc->ptr, s);
    u64 v = 0;
    hip_check(hipMemcpyAsync(&v, acc->ptr, 8, hipMemcpyDeviceToHost, s), "digest");
    hip_check(hipStreamSynchronize(s), "digest");
    return v;
}

PolyPtr poly_uniform(Context& c, u32 level, u32 p_limbs, bool ntt, u64 seed, u64 stream0) {
    PolyPtr p = poly_alloc(c, level, p_limbs, ntt);
    fill_uniform_launch(p->data(), level + 1, p_limbs, c.nq, seed, stream0, c.d_pc, c.log_n, c.stream());
    hip_check(hipGetLastError(), "fill");
    return p;
}

PolyPtr poly_transform(Context& c, const Poly& in, bool forward) {
    if (in.ntt == forward) fail(CKH_ERR_INVALID_PARAMS, forward ? "already in NTT form" : "already in coefficient form");
    use(c, in);
    PolyPtr out = poly_alloc(c, in.level, in.p_limbs, forward);
    c.ntt_poly(in.data(), out->data(), in.limbs(), in.level, in.p_limbs, forward, c.stream());
    hip_check(hipGetLastError(), "ntt");
    return out;
}

PolyPtr poly_automorphism(Context& c, const Poly& in, u64 g) {
    if (g % 2 == 0 || g >= 2 * c.n) fail(CKH_ERR_INVALID_PARAMS, "Galois element must be odd and < 2N");
    use(c, in);
    PolyPtr out = poly_alloc(c, in.level, in.p_limbs, in.ntt);
    c.tick_launch();
    if (in.ntt)
        automorphism_ntt_launch(in.data(), out->data(), in.limbs(), g, c.log_n, c.stream());
    else
        automorphism_coeff_launch(in.data(), out->data(), in.limbs(), in.level, c.nq, g, c.d_pc, c.log_n, c.stream());
    hip_check(hipGetLastError(), "automorphism");
    return out;
}

double time_ntt(Context& c, const std::vector<const Poly*>& polys, bool forward, u32 reps) {
    std::vector<LimbRef> refs;
    for (const Poly* p : polys) {
        use(c, *p);
        for (u32 i = 0; i < p->limbs(); ++i) {
            u32 prime = i <= p->level ? i : c.nq + (i - p->level - 1);
            refs.push_back({p->data() + i * c.n, p->data() + i * c.n, prime, 0});
        }
    }
    if (refs.empty() || reps == 0) fail(CKH_ERR_INVALID_PARAMS, "nothing to time");
    hipStream_t s = c.stream();
    // The descriptors, then one zeroed control block per launch (the one-launch forward transform
    // needs a fresh one each time).
    const u32 nr = (u32)refs.size();
    const u64 ref_bytes = ((u64)nr * sizeof(LimbRef) + 255) & ~(u64)255;
    const u64 ctl_bytes = ((fused_ctl_words(nr) * 4) + 255) & ~(u64)255;
    std::vector<uint8_t> blob(ref_bytes + (reps + 1) * ctl_bytes, 0);
    std::memcpy(blob.data(), refs.data(), nr * sizeof(LimbRef));
    BufPtr d = c.upload_blob(blob.data(), blob.size(), s);
    const LimbRef* dr = (const LimbRef*)d->ptr;
    const bool fused = c.fused_usable();
    auto launch = [&](u32 r) {
        u32* ctl = (u32*)((uint8_t*)d->ptr + ref_bytes + r * ctl_bytes);
        if (fused && forward)
            ntt_fast_forward_fused(c.ntt, dr, nr, nullptr, ctl, c.fused, s);
        else if (fused)
            ntt_fast_inverse_fused(c.ntt, dr, nr, ctl, c.fused, s);
        else
            launch_ntt(c, dr, nr, forward, s);
    };
    launch(reps);  // warm up
    hipEvent_t a, b;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    hip_check(hipEventRecord(a, s), "event");
    for (u32 r = 0; r < reps; ++r) launch(r);
    hip_check(hipEventRecord(b, s), "event");
    hip_check(hipEventSynchronize(b), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, b), "event");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
}

double time_copy(Context& c, u64 bytes, u32 reps) {
    if (bytes < 16 || bytes % 16 || reps == 0) fail(CKH_ERR_INVALID_PARAMS, "copy size must be a positive multiple of 16");
    hipStream_t s = c.stream();
    BufPtr a = c.alloc(bytes / 8, s), b = c.alloc(bytes / 8, s);
    stream_copy_launch(a->ptr, b->ptr, bytes, s);  // warm up
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event");
    for (u32 r = 0; r < reps; ++r) stream_copy_launch(a->ptr, b->ptr, bytes, s);
    hip_check(hipEventRecord(e1, s), "event");
    hip_check(hipEventSynchronize(e1), "event");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
}

PolyPtr poly_rescale(Context& c, const Poly& in) {
    if (!in.ntt || in.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "rescale needs an NTT-form Q polynomial");
    if (in.level == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "cannot rescale at level 0");
    use(c, in);
    PolyPtr out = poly_alloc(c, in.level - 1, 0, true);
    BufPtr scratch = c.alloc((u64)(1 + in.level) * c.n);
    c.rescale({{in.data(), in.level, out->data(), scratch->ptr, scratch->ptr + c.n}}, c.stream());
    hip_check(hipGetLastError(), "rescale");
    return out;
}

// RnsPoly::mul_scalar_assign (crates/ckks-train-core/src/poly.rs:184-189; the scalar reduced mod
// each q_i and applied as a Shoup product), then RnsPoly::neg (poly.rs:157-161) when neg is set.
PolyPtr poly_scalar_mul(Context& c, const Poly& a, u64 k, bool neg) {
    if (a.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "scalar products take Q polynomials");
    use(c, a);
    PolyPtr out = poly_alloc(c, a.level, 0, a.ntt);
    std::vector<u64> ws(2 * (a.level + 1));
    for (u32 i = 0; i <= a.level; ++i) {
        const u64 q = c.primes[i], w = k % q;
        ws[2 * i] = w;
        ws[2 * i + 1] = (u64)(((unsigned __int128)w << 64) / q);
    }
    hipStream_t s = c.stream();
    BufPtr consts = c.upload_blob(ws.data(), ws.size() * 8, s);
    c.tick_launch();
    scalar_mul_launch(a.data(), out->data(), a.level + 1, consts->ptr, neg, c.d_pc, c.log_n, s);
    hip_check(hipGetLastError(), "scalar product");
    return out;
}

// RnsPoly::raise_from_level0 (crates/ckks-train-core/src/poly.rs:298-345).
PolyPtr poly_raise(Context& c, const Poly& a, u32 top) {
    if (a.level != 0 || a.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "raise_from_level0 needs a level-0 Q polynomial");
    if (top >= c.nq) fail(CKH_ERR_INVALID_PARAMS, "level out of range");
    use(c, a);
    PolyPtr out = poly_alloc(c, top, 0, a.ntt);
    BufPtr x = c.alloc(c.n);
    c.raise_level0({{a.data(), top, out->data(), x->ptr, a.ntt}}, c.stream());
    hip_check(hipGetLastError(), "raise");
    return out;
}

PolyPtr poly_pointwise(Context& c, PwOp op, const Poly& a, const Poly* b) {
    if (b && (b->level != a.level || b->p_limbs != a.p_limbs || b->ntt != a.ntt))
        fail(CKH_ERR_LEVEL_MISMATCH, "operands at different levels or forms");
    if (a.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "pointwise ops take Q polynomials");
    use(c, a);
    if (b) use(c, *b);
    PolyPtr out = poly_alloc(c, a.level, 0, a.ntt);
    c.tick_launch();
    pointwise_launch(op, a.data(), b ? b->data() : nullptr, out->data(), a.level + 1, 0, c.d_pc, c.log_n, c.stream());
    hip_check(hipGetLastError(), "pointwise");
    return out;
}

PolyPtr poly_mod_down(Context& c, const Poly& qp) {
    if (!qp.ntt || qp.p_limbs != c.np) fail(CKH_ERR_INVALID_PARAMS, "mod_down needs an NTT-form QpPoly");
    use(c, qp);
    const u64 L = qp.level + 1;
    PolyPtr out = poly_alloc(c, qp.level, 0, true);
    BufPtr scratch = c.alloc((c.np + L) * c.n);
    c.mod_down({{qp.data(), qp.level, out->data(), nullptr, scratch->ptr, scratch->ptr + c.np * c.n}}, c.stream());
    hip_check(hipGetLastError(), "mod_down");
    return out;
}

// ------------------------------------------------------------------ key switching

KskPtr ksk_upload(Context& c, const u64* host, u32 max_level, u64 galois) {
    if (max_level >= c.nq) fail(CKH_ERR_INVALID_PARAMS, "key level above max_level");
    if (galois != 0 && (galois % 2 == 0 || galois >= 2 * c.n))
        fail(CKH_ERR_INVALID_PARAMS, "Galois element must be odd and < 2N");
    auto k = std::make_shared<Ksk>();
    k->ctx = &c;
    k->max_level = max_level;
    k->parts = c.ndigits(max_level);
    k->galois = galois;
    const u64 words = (u64)k->parts * 2 * k->poly_words();
    k->buf = c.alloc(words);
    if (host) {
        c.tick_upload();
        const u32 lp = max_level + 1 + c.np;
        for (u32 pj = 0; pj < 2 * k->parts; ++pj)
            for (u32 i = 0; i < lp; ++i) {
                const u64 q = c.primes[i <= max_level ? i : c.nq + (i - max_level - 1)];
                const u64* row = host + ((u64)pj * lp + i) * c.n;
                for (u64 x = 0; x < c.n; ++x)
                    if (row[x] >= q) fail(CKH_ERR_INVALID_PARAMS, "key entry not reduced");
            }
        hipStream_t s = c.stream();
        c.h2d(k->buf->ptr, host, words * 8, s);
        hip_check(hipStreamSynchronize(s), "key upload");
    }
    return k;
}

// Uniform key in the device order: polynomial j (b_0, a_0, b_1, ...) uses generator streams
// 1000 (j + 1) + limb, so a host copy is reproducible.
KskPtr ksk_uniform(Context& c, u32 max_level, u64 galois, u64 seed) {
    KskPtr k = ksk_upload(c, nullptr, max_level, galois);
    for (u32 j = 0; j < 2 * k->parts; ++j)
        fill_uniform_launch(k->buf->ptr + (u64)j * k->poly_words(), max_level + 1, c.np, c.nq, seed, 1000ull * (j + 1),
                            c.d_pc, c.log_n, c.stream());
    hip_check(hipGetLastError(), "fill");
    return k;
}

namespace {
// Scratch words of one key switch at level l: x (L), digits, acc (2 (L + k)), y (2k), z (2L).
u64 ks_scratch_words(const Context& c, u32 level, bool with_digits) {
    const u64 L = level + 1, k = c.np;
    return c.n * (L + (with_digits ? c.ndigits(level) * (L + k) : 0) + 2 * (L + k) + 2 * k + 2 * L);
}

struct KsReq {
    const Poly* d;
    const Ksk* key;
    const u64* add_b;
    const u64* add_a;
    u64* out_b;
    u64* out_a;
};

void check_switch(const Context& c, const Poly& d, const Ksk& key) {
    if (!d.ntt || d.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "key switching needs an NTT-form Q polynomial");
    if (key.ctx != &c) fail(CKH_ERR_INVALID_PARAMS, "key is not over this context");
    if (d.level > key.max_level) fail(CKH_ERR_INVALID_PARAMS, "key does not serve this level");
    use(c, d);
    use(c, key);
}

// The program of a batch of requests (decompose + apply, one launch per stage), its scratch laid
// out from base (ks_scratch_words per request, in request order).
static Program ks_program(Context& c, const std::vector<KsReq>& reqs, u64* base,
                          const std::vector<AutJob>* pre = nullptr) {
    u64 limbs = 0;
    for (auto& r : reqs) limbs += r.d->level + 1;
    const bool small = limbs <= c.small_limbs;
    const bool fuse = (small ? c.fuse_ip_s : c.fuse_ip) && c.fuse_ip_possible();
    std::vector<DecTask> dec;
    std::vector<ApplyTask> app;
    u64* p = base;
    for (auto& r : reqs) {
        const u64 L = r.d->level + 1, k = c.np;
        u64* x = p;
        u64* dig = x + L * c.n;
        u64* acc = dig + (u64)c.ndigits(r.d->level) * (L + k) * c.n;
        u64* y = acc + 2 * (L + k) * c.n;
        u64* z = y + 2 * k * c.n;
        p = z + 2 * L * c.n;
        dec.push_back({r.d->data(), r.d->level, dig, x, false, fuse});
        app.push_back({dig, r.d->level, 0, r.key, r.out_b, r.out_a, r.add_b, r.add_a, acc, y, z, r.d->data(), fuse});
    }
    Program prog;  // one program (one graph when captured) per batched switch
    if (pre) prog.step(S_AUTOMORPH, prog.add(*pre), (u32)pre->size());
    c.prog_decompose(prog, dec, small ? c.conv_variant_s : c.conv_variant);
    c.prog_apply(prog, app);
    if (small) prog.set_conv_schedule(c.conv_variant_s, c.conv_tg_s);
    return prog;
}

// decompose + apply for a batch of requests, in one launch per stage.
static void run_ks_on(Context& c, const std::vector<KsReq>& reqs, hipStream_t s,
                      const std::vector<AutJob>* pre = nullptr) {
    u64 total = 0;
    for (auto& r : reqs) total += ks_scratch_words(c, r.d->level, true);
    BufPtr scratch = c.alloc(total, s);
    c.execute(ks_program(c, reqs, scratch->ptr, pre), s);
    hip_check(hipGetLastError(), "key switch");
}

void run_ks(Context& c, const std::vector<KsReq>& reqs) { run_ks_on(c, reqs, c.stream()); }
}  // namespace

std::vector<PolyPtr> ks_decompose(Context& c, const Poly& d) {
    if (!d.ntt || d.p_limbs) fail(CKH_ERR_INVALID_PARAMS, "key switching needs an NTT-form Q polynomial");
    use(c, d);
    const u32 nd = c.ndigits(d.level);
    const u64 dw = c.digit_words(d.level);
    BufPtr digs = c.alloc(nd * dw);
    BufPtr x = c.alloc((u64)(d.level + 1) * c.n);
    c.decompose({{d.data(), d.level, digs->ptr, x->ptr}}, c.stream());
    hip_check(hipGetLastError(), "decompose");
    std::vector<PolyPtr> out;
    for (u32 j = 0; j < nd; ++j) out.push_back(view(digs, j * dw, d.level, c.np, true));
    return out;
}

std::pair<PolyPtr, PolyPtr> ks_apply(Context& c, const std::vector<PolyPtr>& digits, const Ksk& key) {
    if (digits.empty()) fail(CKH_ERR_INVALID_PARAMS, "no digits to switch");
    const u32 level = digits[0]->level;
    const u32 nd = c.ndigits(level);
    if (digits.size() != nd) fail(CKH_ERR_INVALID_PARAMS, "digit count does not match the level");
    if (nd > key.parts || level > key.max_level) fail(CKH_ERR_INVALID_PARAMS, "key does not serve this level");
    if (key.ctx != &c) fail(CKH_ERR_INVALID_PARAMS, "key is not over this context");
    const u64 dw = c.digit_words(level);
    bool contiguous = true;
    for (u32 j = 0; j < nd; ++j) {
        const Poly& d = *digits[j];
        if (d.level != level || d.p_limbs != c.np || !d.ntt)
            fail(CKH_ERR_INVALID_PARAMS, "digits must share one level and be NTT-form QpPolys");
        if (d.data() != digits[0]->data() + j * dw) contiguous = false;
        use(c, d);
    }
    use(c, key);
    hipStream_t s = c.stream();
    BufPtr packed;
    const u64* dig = digits[0]->data();
    if (!contiguous) {
        packed = c.alloc(nd * dw);
        for (u32 j = 0; j < nd; ++j)
            hip_check(hipMemcpyAsync(packed->ptr + j * dw, digits[j]->data(), dw * 8, hipMemcpyDeviceToDevice, s),
                      "copy");
        dig = packed->ptr;
    }
    const u64 L = level + 1;
    BufPtr out = c.alloc(2 * L * c.n);
    BufPtr scratch = c.alloc(ks_scratch_words(c, level, false));
    u64* acc = scratch->ptr + L * c.n;
    u64* y = acc + 2 * (L + c.np) * c.n;
    u64* z = y + 2 * c.np * c.n;
    c.apply({{dig, level, 0, &key, out->ptr, out->ptr + L * c.n, nullptr, nullptr, acc, y, z}}, s);
    hip_check(hipGetLastError(), "apply_key");
    return {view(out, 0, level, 0, true), view(out, L * c.n, level, 0, true)};
}

std::pair<PolyPtr, PolyPtr> ks_switch(Context& c, const Poly& d, const Ksk& key) {
    check_switch(c, d, key);
    const u64 L = d.level + 1;
    BufPtr out = c.alloc(2 * L * c.n);
    run_ks(c, {{&d, &key, nullptr, nullptr, out->ptr, out->ptr + L * c.n}});
    return {view(out, 0, d.level, 0, true), view(out, L * c.n, d.level, 0, true)};
}

std::vector<std::pair<PolyPtr, PolyPtr>> ks_switch_batch(Context& c, const std::vector<const Poly*>& ds,
                                                         const Ksk& key) {
    std::vector<KsReq> reqs;
    std::vector<std::pair<PolyPtr, PolyPtr>> res;
    u64 total = 0;
    for (const Poly* d : ds) {
        check_switch(c, *d, key);
        total += 2 * (u64)(d->level + 1) * c.n;
    }
    if (ds.empty()) return res;
    BufPtr out = c.alloc(total);
    u64 off = 0;
    for (const Poly* d : ds) {
        const u64 L = d->level + 1;
        reqs.push_back({d, &key, nullptr, nullptr, out->ptr + off, out->ptr + off + L * c.n});
        res.push_back({view(out, off, d->level, 0, true), view(out, off + L * c.n, d->level, 0, true)});
        off += 2 * L * c.n;
    }
    run_ks(c, reqs);
    return res;
}

void ks_switch_batch_static(Context& c, const std::vector<const Poly*>& ds, const Ksk& key,
                            const std::vector<Poly*>& kb, const std::vector<Poly*>& ka) {
    if (kb.size() != ds.size() || ka.size() != ds.size()) fail(CKH_ERR_INVALID_PARAMS, "one output pair per input");
    // Every word range the call reads or writes: inputs, then outputs; outputs may overlap nothing.
    std::vector<std::pair<const u64*, const u64*>> in_r, out_r;
    u64 words = 0;
    for (size_t i = 0; i < ds.size(); ++i) {
        const Poly& d = *ds[i];
        check_switch(c, d, key);
        for (const Poly* o : {kb[i], ka[i]}) {
            if (o->buf->ctx != &c || o->p_limbs || !o->ntt || o->level != d.level)
                fail(CKH_ERR_INVALID_PARAMS, "outputs must be NTT-form Q polynomials of this context at the input's level");
            out_r.push_back({o->data(), o->data() + (u64)(o->level + 1) * c.n});
        }
        in_r.push_back({d.data(), d.data() + (u64)(d.level + 1) * c.n});
        words += ks_scratch_words(c, d.level, true);
    }
    if (ds.empty()) return;
    std::sort(out_r.begin(), out_r.end());
    for (size_t i = 1; i < out_r.size(); ++i)
        if (out_r[i].first < out_r[i - 1].second) fail(CKH_ERR_INVALID_PARAMS, "outputs overlap each other");
    for (auto& o : out_r)
        for (auto& r : in_r)
            if (o.first < r.second && r.first < o.second) fail(CKH_ERR_INVALID_PARAMS, "an output overlaps an input");
    const u64 scratch_bytes = words * 8;
    if (scratch_bytes > CKH_STATIC_SCRATCH_BYTES) fail(CKH_ERR_INVALID_PARAMS, "batch needs more than the static workspace");
    std::vector<KsReq> reqs;
    for (size_t i = 0; i < ds.size(); ++i) {
        use(c, *kb[i]);
        use(c, *ka[i]);
        reqs.push_back({ds[i], &key, nullptr, nullptr, kb[i]->data(), ka[i]->data()});
    }
    // Workspace: the switches' scratch from the start, the descriptor bytes after it.
    Program prog = ks_program(c, reqs, (u64*)c.static_scratch);
    const u64 desc_off = (scratch_bytes + 255) & ~(u64)255;
    if (desc_off + prog.bytes.size() > CKH_STATIC_SCRATCH_BYTES)
        fail(CKH_ERR_INVALID_PARAMS, "batch needs more than the static workspace");
    hipStream_t s = c.stream();
    c.tick_launch();
    c.h2d(c.static_scratch + desc_off, prog.bytes.data(), prog.bytes.size(), s);
    c.run(prog, c.static_scratch + desc_off, s);
    hip_check(hipGetLastError(), "key switch");
}

u64 galois_elt(u64 n, int32_t step) {
    const u64 m = 2 * n;
    const int64_t slots = (int64_t)std::max<u64>(n / 2, 1);
    int64_t k0 = (int64_t)step % slots;
    if (k0 < 0) k0 += slots;
    u64 k = (u64)k0, base = 3 % m, acc = 1 % m;
    while (k) {
        if (k & 1) acc = acc * base % m;
        base = base * base % m;
        k >>= 1;
    }
    return acc;
}

// ------------------------------------------------------------------ evaluator

Evaluator::Evaluator(Context& c, KskPtr r, std::vector<KskPtr> g) : ctx(c), relin(std::move(r)) {
    for (auto& k : g) galois[k->galois] = k;
}

namespace {
// Counts of a composite operation in progress on this thread (innermost last): its nested
// operations count here, and the total reaches the evaluator only when it succeeds
// (crates/ckks-train-cuda/src/matvec.rs `counted`: a device error part-way leaves the counter
// unchanged).
struct PendingCounts {
    const Evaluator* ev;
    ckh_op_counter c;
};
thread_local std::vector<PendingCounts> tl_pending;

void add_into(ckh_op_counter& a, const ckh_op_counter& b) {
    a.rotations += b.rotations;
    a.relins += b.relins;
    a.ct_ct_tensor += b.ct_ct_tensor;
    a.ct_pt_mults += b.ct_pt_mults;
    a.rescales += b.rescales;
    a.sse_switches += b.sse_switches;
    a.bootstraps += b.bootstraps;
}
}  // namespace

template <class F>
auto Evaluator::counted(F&& body) -> decltype(body()) {
    tl_pending.push_back({this, ckh_op_counter{}});
    try {
        auto r = body();
        const ckh_op_counter c = tl_pending.back().c;
        tl_pending.pop_back();
        for (auto it = tl_pending.rbegin(); it != tl_pending.rend(); ++it)
            if (it->ev == this) {
                add_into(it->c, c);
                return r;
            }
        add_counts(c);
        return r;
    } catch (...) {
        tl_pending.pop_back();
        throw;
    }
}

void Evaluator::count(u64 ckh_op_counter::*field, u64 v) {
    for (auto it = tl_pending.rbegin(); it != tl_pending.rend(); ++it)
        if (it->ev == this) {
            it->c.*field += v;
            return;
        }
    std::lock_guard<std::mutex> lk(mu_);
    cnt_.*field += v;
}

ckh_op_counter Evaluator::counter() {
    std::lock_guard<std::mutex> lk(mu_);
    return cnt_;
}

void Evaluator::reset_counter() {
    std::lock_guard<std::mutex> lk(mu_);
    cnt_ = ckh_op_counter{};
}

void Evaluator::add_counts(const ckh_op_counter& c) {
    std::lock_guard<std::mutex> lk(mu_);
    cnt_.rotations += c.rotations;
    cnt_.relins += c.relins;
    cnt_.ct_ct_tensor += c.ct_ct_tensor;
    cnt_.ct_pt_mults += c.ct_pt_mults;
    cnt_.rescales += c.rescales;
    cnt_.sse_switches += c.sse_switches;
    cnt_.bootstraps += c.bootstraps;
}

bool Evaluator::owns(const Ct& ct) const {
    for (auto& p : ct.comps)
        if (!p || p->buf->ctx != &ctx || p->p_limbs || !p->ntt) return false;
    return !ct.comps.empty();
}

void Evaluator::check_basis(const Ct& ct) const {
    if (!owns(ct)) fail(CKH_ERR_INVALID_PARAMS, "operand is not over this evaluator's data basis");
    for (auto& p : ct.comps) use(ctx, *p);
}

static void check_level(u32 a, u32 b) {
    if (a != b) fail(CKH_ERR_LEVEL_MISMATCH, "level mismatch");
}

// SCALE_TOLERANCE = 2^-10 (evaluator.rs:96).
static void check_scale(double a, double b) {
    if (!(std::fabs(a - b) < (1.0 / 1024.0) * std::min(a, b))) fail(CKH_ERR_SCALE_MISMATCH, "scale mismatch");
}

static CtPtr make_ct(std::vector<PolyPtr> comps, double scale) {
    auto c = std::make_shared<Ct>();
    c->comps = std::move(comps);
    c->scale = scale;
    return c;
}

bool Evaluator::has_galois_key(int32_t step) const {
    const u64 e = rotation_elt(step);
    return e != 0 && galois.count(e);
}

u64 Evaluator::rotation_elt(int32_t steps) const {
    const int64_t slots = (int64_t)(ctx.n / 2);
    int64_t k = (int64_t)steps % slots;
    if (k < 0) k += slots;
    return k == 0 ? 0 : galois_elt(ctx.n, (int32_t)k);
}

const Ksk& Evaluator::relin_key_checked(u32 level) const {
    if (!relin) fail(CKH_ERR_INVALID_PARAMS, "evaluator has no relinearization key");
    if (level > relin->max_level) fail(CKH_ERR_INVALID_PARAMS, "relinearization key does not serve this level");
    use(ctx, *relin);
    return *relin;
}

const Ksk& Evaluator::galois_key(u64 elt, u32 level) const {
    if (elt % 2 == 0 || elt >= 2 * ctx.n) fail(CKH_ERR_INVALID_PARAMS, "Galois element must be odd and < 2N");
    auto it = galois.find(elt);
    if (it == galois.end()) fail(CKH_ERR_INVALID_PARAMS, "missing galois key");
    if (level > it->second->max_level) fail(CKH_ERR_INVALID_PARAMS, "galois key does not serve this level");
    use(ctx, *it->second);
    return *it->second;
}

// Evaluator::add_sub (evaluator.rs:694-716): the shorter operand is padded with zeros.
CtPtr Evaluator::add(const Ct& a, const Ct& b, bool sub) {
    check_basis(a);
    check_basis(b);
    check_level(a.level(), b.level());
    check_scale(a.scale, b.scale);
    std::vector<PolyPtr> out;
    for (size_t i = 0; i < std::max(a.comps.size(), b.comps.size()); ++i) {
        if (i < a.comps.size() && i < b.comps.size())
            out.push_back(poly_pointwise(ctx, sub ? PW_SUB : PW_ADD, *a.comps[i], b.comps[i].get()));
        else if (i < a.comps.size())
            out.push_back(a.comps[i]);
        else
            out.push_back(sub ? poly_pointwise(ctx, PW_NEG, *b.comps[i], nullptr) : b.comps[i]);
    }
    return make_ct(std::move(out), a.scale);
}

CtPtr Evaluator::negate(const Ct& a) {
    check_basis(a);
    std::vector<PolyPtr> out;
    for (auto& p : a.comps) out.push_back(poly_pointwise(ctx, PW_NEG, *p, nullptr));
    return make_ct(std::move(out), a.scale);
}

static void check_plain(const Context& c, const Poly& pt) {
    if (pt.buf->ctx != &c || pt.p_limbs || !pt.ntt)
        fail(CKH_ERR_INVALID_PARAMS, "operand is not over this evaluator's data basis");
    use(c, pt);
}

// Evaluator::add_plain (evaluator.rs:287-296): c0 + p, scales within tolerance.
CtPtr Evaluator::add_plain(const Ct& a, const Poly& pt, double pt_scale) {
    check_basis(a);
    check_plain(ctx, pt);
    check_level(a.level(), pt.level);
    check_scale(a.scale, pt_scale);
    std::vector<PolyPtr> out = a.comps;
    out[0] = poly_pointwise(ctx, PW_ADD, *a.comps[0], &pt);
    return make_ct(std::move(out), a.scale);
}

// Evaluator::mul_plain (evaluator.rs:298-308): every component times p, scale product.
CtPtr Evaluator::mul_plain(const Ct& a, const Poly& pt, double pt_scale) {
    check_basis(a);
    check_plain(ctx, pt);
    check_level(a.level(), pt.level);
    std::vector<PolyPtr> out;
    for (auto& c : a.comps) out.push_back(poly_pointwise(ctx, PW_MUL, *c, &pt));
    CtPtr r = make_ct(std::move(out), a.scale * pt_scale);
    count(&ckh_op_counter::ct_pt_mults);
    return r;
}

void Evaluator::check_mul_operands(const Ct& a, const Ct& b) const {
    check_basis(a);
    check_basis(b);
    if (a.comps.size() != 2 || b.comps.size() != 2)
        fail(CKH_ERR_INVALID_PARAMS, "multiplication needs 2-component ciphertexts; relinearize first");
    check_level(a.level(), b.level());
    if (a.level() == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "no levels left");
}

// Evaluator::mul_no_relin (evaluator.rs:422-438).
CtPtr Evaluator::mul_no_relin(const Ct& a, const Ct& b) {
    check_mul_operands(a, b);
    auto& x = a.comps;
    auto& y = b.comps;
    PolyPtr d0 = poly_pointwise(ctx, PW_MUL, *x[0], y[0].get());
    PolyPtr t0 = poly_pointwise(ctx, PW_MUL, *x[0], y[1].get());
    PolyPtr t1 = poly_pointwise(ctx, PW_MUL, *x[1], y[0].get());
    PolyPtr d1 = poly_pointwise(ctx, PW_ADD, *t0, t1.get());
    PolyPtr d2 = poly_pointwise(ctx, PW_MUL, *x[1], y[1].get());
    CtPtr r = make_ct({d0, d1, d2}, a.scale * b.scale);
    count(&ckh_op_counter::ct_ct_tensor);
    return r;
}

// mul_no_relin of several pairs in one launch: per output limb the product (two products summed
// for the middle component) as a multiply-sum job; the same residues as mul_no_relin.
std::vector<CtPtr> Evaluator::mul_no_relin_batch(const std::vector<std::pair<const Ct*, const Ct*>>& pairs) {
    u64 total = 0;
    for (const auto& pr : pairs) {
        check_mul_operands(*pr.first, *pr.second);
        total += 3 * (u64)(pr.first->level() + 1) * ctx.n;
    }
    std::vector<CtPtr> res;
    if (pairs.empty()) return res;
    BufPtr out = ctx.alloc(total);
    std::vector<MulSumJob> jobs;
    std::vector<MulSumTerm> terms;
    u64 off = 0;
    for (const auto& pr : pairs) {
        const u32 level = pr.first->level();
        const u64 L = level + 1, n = ctx.n;
        const auto& x = pr.first->comps;
        const auto& y = pr.second->comps;
        for (const Poly* q : {x[0].get(), x[1].get(), y[0].get(), y[1].get()}) use(ctx, *q);
        u64* d0 = out->ptr + off;
        u64* d1 = d0 + L * n;
        u64* d2 = d1 + L * n;
        for (u32 i = 0; i < L; ++i) {
            const u64 o = (u64)i * n;
            jobs.push_back({d0 + o, (u32)terms.size(), 1, i, 0});
            terms.push_back({x[0]->data() + o, y[0]->data() + o});
            jobs.push_back({d1 + o, (u32)terms.size(), 2, i, 0});
            terms.push_back({x[0]->data() + o, y[1]->data() + o});
            terms.push_back({x[1]->data() + o, y[0]->data() + o});
            jobs.push_back({d2 + o, (u32)terms.size(), 1, i, 0});
            terms.push_back({x[1]->data() + o, y[1]->data() + o});
        }
        res.push_back(make_ct({view(out, off, level, 0, true), view(out, off + L * n, level, 0, true),
                               view(out, off + 2 * L * n, level, 0, true)},
                              pr.first->scale * pr.second->scale));
        off += 3 * L * n;
    }
    Program prog;
    const size_t oj = prog.add(jobs), ot = prog.add(terms);
    prog.step(S_MUL_SUM, oj, (u32)jobs.size(), ot);
    ctx.execute(prog, ctx.stream());
    hip_check(hipGetLastError(), "tensor products");
    count(&ckh_op_counter::ct_ct_tensor, pairs.size());
    return res;
}

// Evaluator::relinearize (evaluator.rs:441-458): (c0 + kb, c1 + ka), the additions fused into
// the mod_down finish.
CtPtr Evaluator::relinearize(const Ct& ct) {
    check_basis(ct);
    if (ct.comps.size() != 3) fail(CKH_ERR_INVALID_PARAMS, "relinearize needs a 3-component ciphertext");
    const Ksk& key = relin_key_checked(ct.level());
    const u32 level = ct.level();
    const u64 L = level + 1;
    BufPtr out = ctx.alloc(2 * L * ctx.n);
    run_ks(ctx, {{ct.comps[2].get(), &key, ct.comps[0]->data(), ct.comps[1]->data(), out->ptr, out->ptr + L * ctx.n}});
    CtPtr r = make_ct({view(out, 0, level, 0, true), view(out, L * ctx.n, level, 0, true)}, ct.scale);
    count(&ckh_op_counter::relins);
    return r;
}

std::vector<CtPtr> Evaluator::relinearize_batch(const std::vector<const Ct*>& cts) {
    u64 total = 0;
    for (const Ct* ct : cts) {
        check_basis(*ct);
        if (ct->comps.size() != 3) fail(CKH_ERR_INVALID_PARAMS, "relinearize needs a 3-component ciphertext");
        relin_key_checked(ct->level());
        total += 2 * (u64)(ct->level() + 1) * ctx.n;
    }
    std::vector<CtPtr> res;
    if (cts.empty()) return res;
    BufPtr out = ctx.alloc(total);
    std::vector<KsReq> reqs;
    u64 off = 0;
    for (const Ct* ct : cts) {
        const u32 level = ct->level();
        const u64 L = level + 1;
        reqs.push_back({ct->comps[2].get(), relin.get(), ct->comps[0]->data(), ct->comps[1]->data(), out->ptr + off,
                        out->ptr + off + L * ctx.n});
        res.push_back(make_ct({view(out, off, level, 0, true), view(out, off + L * ctx.n, level, 0, true)}, ct->scale));
        off += 2 * L * ctx.n;
    }
    run_ks(ctx, reqs);
    count(&ckh_op_counter::relins, cts.size());
    return res;
}

CtPtr Evaluator::mul(const Ct& a, const Ct& b) {
    check_mul_operands(a, b);
    relin_key_checked(a.level());
    return counted([&] { return relinearize(*mul_no_relin(a, b)); });
}

// Evaluator::rescale (evaluator.rs:498-509): scale / q_l as f64.
CtPtr Evaluator::rescale(const Ct& ct) {
    check_basis(ct);
    const u32 l = ct.level();
    if (l == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "no levels left");
    const u64 L = l;  // output limbs
    BufPtr out = ctx.alloc((u64)ct.comps.size() * L * ctx.n);
    BufPtr scratch = ctx.alloc((u64)ct.comps.size() * (1 + L) * ctx.n);
    std::vector<RescaleTask> tasks;
    std::vector<PolyPtr> comps;
    for (size_t i = 0; i < ct.comps.size(); ++i) {
        u64* x = scratch->ptr + i * (1 + L) * ctx.n;
        tasks.push_back({ct.comps[i]->data(), l, out->ptr + i * L * ctx.n, x, x + ctx.n});
        comps.push_back(view(out, i * L * ctx.n, l - 1, 0, true));
    }
    ctx.rescale(tasks, ctx.stream());
    hip_check(hipGetLastError(), "rescale");
    CtPtr r = make_ct(std::move(comps), ct.scale / (double)ctx.primes[l]);
    count(&ckh_op_counter::rescales);
    return r;
}

std::vector<CtPtr> Evaluator::rescale_batch(const std::vector<const Ct*>& cts) {
    u64 out_words = 0, scratch_words = 0;
    for (const Ct* ct : cts) {
        check_basis(*ct);
        const u32 l = ct->level();
        if (l == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "no levels left");
        out_words += (u64)ct->comps.size() * l * ctx.n;
        scratch_words += (u64)ct->comps.size() * (1 + l) * ctx.n;
    }
    std::vector<CtPtr> res;
    if (cts.empty()) return res;
    BufPtr out = ctx.alloc(out_words);
    BufPtr scratch = ctx.alloc(scratch_words);
    std::vector<RescaleTask> tasks;
    u64 oo = 0, so = 0;
    for (const Ct* ct : cts) {
        const u32 l = ct->level();
        const u64 L = l;  // output limbs
        std::vector<PolyPtr> comps;
        for (const auto& comp : ct->comps) {
            use(ctx, *comp);
            u64* x = scratch->ptr + so;
            tasks.push_back({comp->data(), l, out->ptr + oo, x, x + ctx.n});
            comps.push_back(view(out, oo, l - 1, 0, true));
            oo += L * ctx.n;
            so += (1 + L) * ctx.n;
        }
        res.push_back(make_ct(std::move(comps), ct->scale / (double)ctx.primes[l]));
    }
    ctx.rescale(tasks, ctx.stream());
    hip_check(hipGetLastError(), "rescale");
    count(&ckh_op_counter::rescales, cts.size());
    return res;
}

CtPtr Evaluator::mod_switch_to_level(const Ct& ct, u32 level) {
    check_basis(ct);
    if (level > ct.level()) fail(CKH_ERR_INVALID_PARAMS, "cannot switch a ciphertext up to a higher level");
    std::vector<PolyPtr> comps;
    for (auto& p : ct.comps) comps.push_back(view(p->buf, p->off, level, 0, true));
    return make_ct(std::move(comps), ct.scale);
}

CtPtr Evaluator::relabel(const Ct& ct, double scale) {
    if (!(std::isfinite(scale) && scale > 0)) fail(CKH_ERR_INVALID_PARAMS, "scale must be positive and finite");
    return make_ct(ct.comps, scale);
}

// Evaluator::apply_galois (evaluator.rs:640-647) + finish_galois (800-804).
CtPtr Evaluator::apply_galois(const Ct& ct, u64 elt) {
    const Ksk& key = galois_key(elt, ct.level());
    const u32 level = ct.level();
    const u64 L = level + 1;
    PolyPtr d = poly_automorphism(ctx, *ct.comps[1], elt);
    PolyPtr r0 = poly_automorphism(ctx, *ct.comps[0], elt);
    BufPtr out = ctx.alloc(2 * L * ctx.n);
    run_ks(ctx, {{d.get(), &key, r0->data(), nullptr, out->ptr, out->ptr + L * ctx.n}});
    CtPtr r = make_ct({view(out, 0, level, 0, true), view(out, L * ctx.n, level, 0, true)}, ct.scale);
    count(&ckh_op_counter::rotations);
    return r;
}

static void check_galois_input(const Ct& ct) {
    if (ct.comps.size() != 2)
        fail(CKH_ERR_INVALID_PARAMS, "rotation needs a 2-component ciphertext; relinearize first");
}

CtPtr Evaluator::rotate(const Ct& ct, int32_t steps) {
    check_basis(ct);
    check_galois_input(ct);
    const u64 elt = rotation_elt(steps);
    if (elt == 0) return make_ct(ct.comps, ct.scale);
    return apply_galois(ct, elt);
}

CtPtr Evaluator::conjugate(const Ct& ct) {
    check_basis(ct);
    check_galois_input(ct);
    return apply_galois(ct, 2 * ctx.n - 1);
}

// Evaluator::hoisted_rotate (evaluator.rs:572-611): every key looked up first, one
// decomposition of c1, then every step's inner product (digits read through sigma_g) and
// mod_down in one batch; sigma_g(c0) is added in the mod_down finish.
std::vector<CtPtr> Evaluator::hoisted_rotate(const Ct& ct, const std::vector<int32_t>& steps) {
    check_basis(ct);
    check_galois_input(ct);
    const u32 level = ct.level();
    std::vector<std::pair<u64, const Ksk*>> keys;
    for (int32_t s : steps) {
        const u64 e = rotation_elt(s);
        keys.push_back({e, e ? &galois_key(e, level) : nullptr});
    }
    std::vector<CtPtr> res(steps.size());
    u32 nontrivial = 0;
    for (auto& k : keys) nontrivial += k.first != 0;
    if (nontrivial == 0) {
        for (auto& r : res) r = make_ct(ct.comps, ct.scale);
        return res;
    }
    if (!config.hoist) {
        // naive form: every step an individual rotation (same words, same count)
        return counted([&] {
            for (size_t i = 0; i < steps.size(); ++i)
                res[i] = keys[i].first ? apply_galois(ct, keys[i].first) : make_ct(ct.comps, ct.scale);
            return res;
        });
    }
    const u64 L = level + 1, n = ctx.n, k = ctx.np;
    const u32 nd = ctx.ndigits(level);
    const u64 dw = ctx.digit_words(level);
    hipStream_t s = ctx.stream();
    BufPtr digs = ctx.alloc(nd * dw + L * n);
    ctx.decompose({{ct.comps[1]->data(), level, digs->ptr, digs->ptr + nd * dw}}, s);
    const u64 per = (2 * (L + k) + 2 * k + 2 * L + L) * n;  // acc, y, z, sigma(c0)
    BufPtr scratch = ctx.alloc(per * nontrivial);
    BufPtr out = ctx.alloc(2 * L * n * nontrivial);
    std::vector<ApplyTask> tasks;
    std::vector<AutJob> aut;  // sigma_g(c0) of every step, in the same program
    u32 t = 0;
    for (size_t i = 0; i < steps.size(); ++i) {
        if (!keys[i].first) {
            res[i] = make_ct(ct.comps, ct.scale);
            continue;
        }
        u64* acc = scratch->ptr + t * per;
        u64* y = acc + 2 * (L + k) * n;
        u64* z = y + 2 * k * n;
        u64* r0 = z + 2 * L * n;
        for (u32 li = 0; li < L; ++li) aut.push_back({ct.comps[0]->data() + li * n, r0 + li * n, keys[i].first});
        u64* ob = out->ptr + 2 * L * n * t;
        tasks.push_back({digs->ptr, level, keys[i].first, keys[i].second, ob, ob + L * n, r0, nullptr, acc, y, z});
        res[i] = make_ct({view(out, 2 * L * n * t, level, 0, true), view(out, 2 * L * n * t + L * n, level, 0, true)},
                         ct.scale);
        ++t;
    }
    Program prog;
    prog.step(S_AUTOMORPH, prog.add(aut), (u32)aut.size());
    ctx.prog_apply(prog, tasks);
    ctx.execute(prog, s);
    hip_check(hipGetLastError(), "hoisted rotate");
    count(&ckh_op_counter::rotations, nontrivial);
    return res;
}

// Hoisted rotations of several ciphertexts (each with its own steps) in batched launches: every
// check first (nothing runs if one fails), one decomposition program for all inputs, then one
// program of every step's automorphism, inner product and mod_down. The same words and counts
// as hoisted_rotate per item.
std::vector<std::vector<CtPtr>> Evaluator::hoisted_rotate_batch(
    const std::vector<std::pair<const Ct*, std::vector<int32_t>>>& items) {
    std::vector<std::vector<CtPtr>> res(items.size());
    std::vector<std::vector<std::pair<u64, const Ksk*>>> keys(items.size());
    u64 dig_words = 0, scr_words = 0, out_words = 0;
    u32 nontrivial = 0;
    const u64 n = ctx.n, k = ctx.np;
    for (size_t i = 0; i < items.size(); ++i) {
        const Ct& ct = *items[i].first;
        check_basis(ct);
        check_galois_input(ct);
        const u32 level = ct.level();
        const u64 L = level + 1;
        u32 nt = 0;
        for (int32_t st : items[i].second) {
            const u64 e = rotation_elt(st);
            keys[i].push_back({e, e ? &galois_key(e, level) : nullptr});
            nt += e != 0;
        }
        if (nt) dig_words += ctx.ndigits(level) * ctx.digit_words(level) + L * n;
        scr_words += nt * (2 * (L + k) + 2 * k + 2 * L + L) * n;
        out_words += nt * 2 * L * n;
        nontrivial += nt;
    }
    if (!config.hoist) {
        return counted([&] {
            for (size_t i = 0; i < items.size(); ++i) res[i] = hoisted_rotate(*items[i].first, items[i].second);
            return res;
        });
    }
    if (nontrivial == 0) {
        for (size_t i = 0; i < items.size(); ++i)
            for (size_t j = 0; j < items[i].second.size(); ++j)
                res[i].push_back(make_ct(items[i].first->comps, items[i].first->scale));
        return res;
    }
    hipStream_t s = ctx.stream();
    BufPtr digs = ctx.alloc(dig_words), scratch = ctx.alloc(scr_words), out = ctx.alloc(out_words);
    std::vector<DecTask> dec;
    std::vector<ApplyTask> tasks;
    std::vector<AutJob> aut;
    u64 dpos = 0, spos = 0, opos = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        const Ct& ct = *items[i].first;
        const u32 level = ct.level();
        const u64 L = level + 1;
        const u64 per = (2 * (L + k) + 2 * k + 2 * L + L) * n;
        bool any = false;
        for (auto& kk : keys[i]) any |= kk.first != 0;
        u64* dig = nullptr;
        if (any) {
            dig = digs->ptr + dpos;
            const u64 dlen = ctx.ndigits(level) * ctx.digit_words(level);
            dec.push_back({ct.comps[1]->data(), level, dig, dig + dlen});
            dpos += dlen + L * n;
        }
        for (size_t j = 0; j < keys[i].size(); ++j) {
            const u64 e = keys[i][j].first;
            if (!e) {
                res[i].push_back(make_ct(ct.comps, ct.scale));
                continue;
            }
            u64* acc = scratch->ptr + spos;
            u64* y = acc + 2 * (L + k) * n;
            u64* z = y + 2 * k * n;
            u64* r0 = z + 2 * L * n;
            spos += per;
            for (u32 li = 0; li < L; ++li) aut.push_back({ct.comps[0]->data() + li * n, r0 + li * n, e});
            u64* ob = out->ptr + opos;
            tasks.push_back({dig, level, e, keys[i][j].second, ob, ob + L * n, r0, nullptr, acc, y, z});
            res[i].push_back(make_ct({view(out, opos, level, 0, true), view(out, opos + L * n, level, 0, true)}, ct.scale));
            opos += 2 * L * n;
        }
    }
    ctx.decompose(dec, s);
    Program prog;
    prog.step(S_AUTOMORPH, prog.add(aut), (u32)aut.size());
    ctx.prog_apply(prog, tasks);
    ctx.execute(prog, s);
    hip_check(hipGetLastError(), "hoisted rotate");
    count(&ckh_op_counter::rotations, nontrivial);
    return res;
}

std::vector<CtPtr> Evaluator::rotate_batch(const std::vector<std::pair<const Ct*, int32_t>>& items) {
    std::vector<CtPtr> res(items.size());
    std::vector<std::tuple<size_t, u64, const Ksk*>> todo;
    for (size_t i = 0; i < items.size(); ++i) {
        const Ct& ct = *items[i].first;
        check_basis(ct);
        check_galois_input(ct);
        const u64 e = rotation_elt(items[i].second);
        if (e == 0)
            res[i] = make_ct(ct.comps, ct.scale);
        else
            todo.push_back({i, e, &galois_key(e, ct.level())});
    }
    if (todo.empty()) return res;
    u64 total = 0;
    for (auto& [i, e, k] : todo) total += 4 * (u64)(items[i].first->level() + 1) * ctx.n;
    BufPtr buf = ctx.alloc(total);  // per item: sigma(c1), sigma(c0), kb + sigma(c0), ka
    std::vector<KsReq> reqs;
    std::vector<Poly> dpolys;
    dpolys.reserve(todo.size());
    u64 off = 0;
    std::vector<AutJob> aut;  // sigma_g of both components, in the switch's program
    for (auto& [i, e, k] : todo) {
        const Ct& ct = *items[i].first;
        const u32 level = ct.level();
        const u64 Lw = (u64)(level + 1) * ctx.n;
        u64* d = buf->ptr + off;
        for (u32 li = 0; li <= level; ++li) {
            aut.push_back({ct.comps[1]->data() + li * ctx.n, d + li * ctx.n, e});
            aut.push_back({ct.comps[0]->data() + li * ctx.n, d + Lw + li * ctx.n, e});
        }
        dpolys.push_back(Poly{buf, off, level, 0, true});
        reqs.push_back({&dpolys.back(), k, d + Lw, nullptr, d + 2 * Lw, d + 3 * Lw});
        res[i] = make_ct({view(buf, off + 2 * Lw, level, 0, true), view(buf, off + 3 * Lw, level, 0, true)}, ct.scale);
        off += 4 * Lw;
    }
    run_ks_on(ctx, reqs, ctx.stream(), &aut);
    count(&ckh_op_counter::rotations, todo.size());
    return res;
}

// Evaluator::bsgs for one plaintext-diagonal term (crates/ckks-train-ckks/src/linalg.rs:800-888):
// every check before work, hoisted baby rotations for the baby steps a non-zero entry uses,
// per giant block the sum of mul_plain(rot(x, k2 s), entry) in k2 order, the nonzero giant
// shifts rotated (in one batch), blocks summed in k1 order.
// check_bsgs and check_stride / check_offset (linalg.rs:523-541,677-684): d * stride divides the
// slot count, -d < offset <= 0, and every rotation the product will make has its key (the baby
// steps of the present entries and the nonzero giant shifts of the blocks that hold one).
void Evaluator::check_bsgs_shape(u32 level, u32 baby, u32 giant, u32 stride, int32_t offset,
                                 const std::vector<bool>& present) const {
    if (baby == 0 || giant == 0 || present.size() != (size_t)baby * giant)
        fail(CKH_ERR_INVALID_PARAMS, "entries must hold baby * giant diagonals");
    const u64 slots = ctx.n / 2, d = (u64)baby * giant;
    if (stride == 0 || d * stride > slots || slots % (d * stride) != 0)
        fail(CKH_ERR_INVALID_PARAMS, "d * stride must divide the slot count");
    if (offset > 0 || (int64_t)offset <= -(int64_t)d)
        fail(CKH_ERR_INVALID_PARAMS, "diagonal offset outside -d < offset <= 0");
    bool any = false;
    std::vector<bool> used(baby, false), block(giant, false);
    for (u32 k1 = 0; k1 < giant; ++k1)
        for (u32 k2 = 0; k2 < baby; ++k2)
            if (present[k1 * baby + k2]) used[k2] = block[k1] = any = true;
    if (!any) fail(CKH_ERR_INVALID_PARAMS, "every diagonal entry is zero");
    for (u32 k2 = 1; k2 < baby; ++k2)
        if (used[k2] && rotation_elt((int32_t)(k2 * stride))) galois_key(rotation_elt((int32_t)(k2 * stride)), level);
    for (u32 k1 = 0; k1 < giant; ++k1) {
        const int32_t shift = (offset + (int32_t)(baby * k1)) * (int32_t)stride;
        if (block[k1] && rotation_elt(shift)) galois_key(rotation_elt(shift), level);
    }
}

// The sum of ciphertexts of one level, component count and scale (add applied in turn: the same
// words), every limb of every component in one launch.
CtPtr Evaluator::sum_cts(const std::vector<CtPtr>& cts) {
    if (cts.size() == 1) return cts[0];
    const Ct& c0 = *cts[0];
    for (auto& c : cts) {
        check_basis(*c);
        check_level(c->level(), c0.level());
        check_scale(c->scale, c0.scale);
        if (c->comps.size() != c0.comps.size()) fail(CKH_ERR_INVALID_PARAMS, "sums need one component count");
    }
    const u32 level = c0.level();
    const u64 L = level + 1, n = ctx.n, nc = c0.comps.size();
    BufPtr out = ctx.alloc(nc * L * n);
    std::vector<MulSumJob> jobs;
    std::vector<const u64*> xs;
    for (u64 comp = 0; comp < nc; ++comp)
        for (u32 i = 0; i < L; ++i) {
            jobs.push_back({out->ptr + (comp * L + i) * n, (u32)xs.size(), (u32)cts.size(), i, 0});
            for (auto& c : cts) xs.push_back(c->comps[comp]->data() + i * n);
        }
    Program prog;
    const size_t oj = prog.add(jobs), ox = prog.add(xs);
    prog.step(S_ADD_SUM, oj, (u32)jobs.size(), ox);
    ctx.execute(prog, ctx.stream());
    hip_check(hipGetLastError(), "sum");
    std::vector<PolyPtr> comps;
    for (u64 comp = 0; comp < nc; ++comp) comps.push_back(view(out, comp * L * n, level, 0, true));
    return make_ct(std::move(comps), c0.scale);
}

// xs[t][k2] = rot(x_t, k2 stride) for the baby steps a present entry of term t uses (one hoisted
// call per term).
std::vector<std::vector<CtPtr>> Evaluator::baby_rotations(const std::vector<BsgsTerm>& terms, u32 baby, u32 giant,
                                                          u32 stride) {
    std::vector<std::vector<CtPtr>> xs(terms.size(), std::vector<CtPtr>(baby));
    for (size_t t = 0; t < terms.size(); ++t) {
        std::vector<int32_t> steps;
        std::vector<u32> used_k2;
        for (u32 k2 = 0; k2 < baby; ++k2) {
            bool used = false;
            for (u32 k1 = 0; k1 < giant && !used; ++k1) used = terms[t].present[k1 * baby + k2];
            if (used) {
                steps.push_back((int32_t)(k2 * stride));
                used_k2.push_back(k2);
            }
        }
        std::vector<CtPtr> rot = hoisted_rotate(*terms[t].x, steps);
        for (size_t i = 0; i < used_k2.size(); ++i) xs[t][used_k2[i]] = rot[i];
    }
    return xs;
}

// Plaintext BSGS products with shared giant rotations: every block sum
// sum_t sum_k2 entry_{t,k1,k2} * rot(x_t, k2 stride), both components, in one launch of the
// fused product-sum kernel (the same words as mul_plain and add, which they replace; counted as
// one ct_pt_mults per product), then the blocks' giant rotations in one batched switch.
CtPtr Evaluator::plain_bsgs(const std::vector<BsgsTerm>& terms, const std::vector<std::vector<const Poly*>>& entries,
                            double pt_scale, u32 baby, u32 giant, u32 stride, int32_t offset) {
    std::vector<std::vector<CtPtr>> xs = baby_rotations(terms, baby, giant, stride);
    const Ct& x0 = *terms[0].x;
    const u32 level = x0.level();
    const u64 L = level + 1, n = ctx.n;
    std::vector<u32> blocks;
    u64 products = 0;
    for (u32 k1 = 0; k1 < giant; ++k1) {
        bool any = false;
        for (size_t t = 0; t < terms.size(); ++t)
            for (u32 k2 = 0; k2 < baby; ++k2)
                if (terms[t].present[k1 * baby + k2]) {
                    any = true;
                    ++products;
                }
        if (any) blocks.push_back(k1);
    }
    BufPtr out = ctx.alloc(blocks.size() * 2 * L * n);
    std::vector<MulSumJob> jobs;
    std::vector<MulSumTerm> mterms;
    for (size_t b = 0; b < blocks.size(); ++b) {
        const u32 k1 = blocks[b];
        for (u32 c = 0; c < 2; ++c)
            for (u32 i = 0; i < L; ++i) {
                MulSumJob J{out->ptr + ((2 * b + c) * L + i) * n, (u32)mterms.size(), 0, i, 0};
                for (size_t t = 0; t < terms.size(); ++t)
                    for (u32 k2 = 0; k2 < baby; ++k2) {
                        const size_t e = (size_t)k1 * baby + k2;
                        if (!terms[t].present[e]) continue;
                        mterms.push_back({xs[t][k2]->comps[c]->data() + i * n, entries[t][e]->data() + i * n});
                        ++J.count;
                    }
                jobs.push_back(J);
            }
    }
    Program prog;
    const size_t oj = prog.add(jobs), ot = prog.add(mterms);
    prog.step(S_MUL_SUM, oj, (u32)jobs.size(), ot);
    ctx.execute(prog, ctx.stream());
    hip_check(hipGetLastError(), "bsgs block sums");
    count(&ckh_op_counter::ct_pt_mults, products);
    std::vector<CtPtr> sums;
    std::vector<std::pair<const Ct*, int32_t>> to_rotate;
    for (size_t b = 0; b < blocks.size(); ++b) {
        sums.push_back(make_ct({view(out, 2 * b * L * n, level, 0, true), view(out, (2 * b + 1) * L * n, level, 0, true)},
                               x0.scale * pt_scale));
        to_rotate.push_back({sums.back().get(), (offset + (int32_t)(baby * blocks[b])) * (int32_t)stride});
    }
    return sum_cts(rotate_batch(to_rotate));
}

CtPtr Evaluator::bsgs(const std::vector<BsgsTerm>& terms, u32 baby, u32 giant, u32 stride, int32_t offset,
                      const std::function<CtPtr(size_t, size_t, const Ct&)>& mul,
                      const std::function<CtPtr(const CtPtr&)>& finish) {
    std::vector<std::vector<CtPtr>> xs = baby_rotations(terms, baby, giant, stride);
    // Per block: with bsgs, the products of every term summed, finished once and rotated by the
    // block's giant shift (all blocks' giant rotations in one batched switch); naive, every
    // product finished and rotated by itself. Sums are exact, so their order does not change
    // the words; a block without a present entry costs nothing.
    std::vector<CtPtr> parts;
    std::vector<std::pair<const Ct*, int32_t>> to_rotate;
    for (u32 k1 = 0; k1 < giant; ++k1) {
        const int32_t shift = (offset + (int32_t)(baby * k1)) * (int32_t)stride;
        if (!config.bsgs && shift != 0) {
            for (size_t t = 0; t < terms.size(); ++t)
                for (u32 k2 = 0; k2 < baby; ++k2) {
                    const size_t i = (size_t)k1 * baby + k2;
                    if (!terms[t].present[i]) continue;
                    parts.push_back(finish(mul(t, i, *xs[t][k2])));
                    to_rotate.push_back({parts.back().get(), shift});
                }
            continue;
        }
        CtPtr inner;
        for (size_t t = 0; t < terms.size(); ++t)
            for (u32 k2 = 0; k2 < baby; ++k2) {
                const size_t i = (size_t)k1 * baby + k2;
                if (!terms[t].present[i]) continue;
                CtPtr v = mul(t, i, *xs[t][k2]);
                inner = inner ? add(*inner, *v, false) : v;
            }
        if (!inner) continue;
        parts.push_back(finish(inner));
        to_rotate.push_back({parts.back().get(), shift});
    }
    std::vector<CtPtr> rotated = rotate_batch(to_rotate);
    CtPtr acc;
    for (auto& r : rotated) acc = acc ? add(*acc, *r, false) : r;
    return acc;
}

// Evaluator::matvec_plain_sum_sparse (crates/ckks-train-ckks/src/linalg.rs:694-790): sum_t W_t x_t
// over terms of one shape with the giant rotations shared; NULL entries are skipped (with the
// baby rotations and giant blocks only they would use). Checks in the reference's order, all
// before any work; every rotation step of the shape needs its key.
CtPtr Evaluator::matvec_plain_sum(const std::vector<const Ct*>& xs, const std::vector<std::vector<const Poly*>>& entries,
                                  double pt_scale, u32 baby, u32 giant, u32 stride, int32_t offset) {
    if (xs.empty()) fail(CKH_ERR_INVALID_PARAMS, "matvec_plain_sum needs at least one term");
    if (entries.size() != xs.size()) fail(CKH_ERR_INVALID_PARAMS, "one diagonal set per term");
    std::vector<BsgsTerm> terms;
    for (size_t t = 0; t < xs.size(); ++t) {
        const Ct& x = *xs[t];
        check_basis(x);
        check_galois_input(x);
        if (entries[t].size() != (size_t)baby * giant)
            fail(CKH_ERR_INVALID_PARAMS, "entries must hold baby * giant diagonals");
        BsgsTerm term{&x, std::vector<bool>(entries[t].size())};
        bool any = false;
        for (size_t i = 0; i < entries[t].size(); ++i)
            if (entries[t][i]) term.present[i] = any = true;
        if (!any) fail(CKH_ERR_INVALID_PARAMS, "a term of matvec_plain_sum has no nonzero diagonal");
        for (const Poly* w : entries[t]) {
            if (!w) continue;
            check_plain(ctx, *w);
            check_level(x.level(), w->level);
        }
        check_level(x.level(), xs[0]->level());
        if (x.scale != xs[0]->scale) fail(CKH_ERR_SCALE_MISMATCH, "terms differ in scale");
        std::vector<bool> all(entries[t].size(), true);
        check_bsgs_shape(x.level(), baby, giant, stride, offset, all);
        terms.push_back(std::move(term));
    }
    return counted([&] {
        if (config.bsgs) return plain_bsgs(terms, entries, pt_scale, baby, giant, stride, offset);
        return bsgs(
            terms, baby, giant, stride, offset,
            [&](size_t t, size_t i, const Ct& xr) { return mul_plain(xr, *entries[t][i], pt_scale); },
            [](const CtPtr& c) { return c; });
    });
}

// Evaluator::matvec_plain (crates/ckks-train-ckks/src/linalg.rs:608-621): entries[k1 * baby +
// k2] = rot(diag_{offset + baby k1 + k2}, -(offset + baby k1) stride) (NULL: a zero entry,
// skipped), all at x's level with scale pt_scale.
CtPtr Evaluator::matvec_plain(const Ct& x, const std::vector<const Poly*>& entries, double pt_scale, u32 baby,
                              u32 giant, u32 stride, int32_t offset) {
    check_basis(x);
    check_galois_input(x);
    std::vector<bool> present(entries.size());
    for (size_t i = 0; i < entries.size(); ++i) {
        if (!entries[i]) continue;
        check_plain(ctx, *entries[i]);
        check_level(x.level(), entries[i]->level);
        present[i] = true;
    }
    check_bsgs_shape(x.level(), baby, giant, stride, offset, present);
    return counted([&] {
        if (config.bsgs) return plain_bsgs({{&x, present}}, {entries}, pt_scale, baby, giant, stride, offset);
        return bsgs(
            {{&x, present}}, baby, giant, stride, offset,
            [&](size_t, size_t i, const Ct& xr) { return mul_plain(xr, *entries[i], pt_scale); },
            [](const CtPtr& c) { return c; });
    });
}

// Evaluator::matvec_enc (linalg.rs:638-665): encrypted 2-component diagonals; lazy_relin: per
// block the tensor products summed at degree 2 and relinearized once, else one mul per product.
// The checks of matvec_enc, before any work; returns which entries are present.
std::vector<bool> Evaluator::check_matvec_enc(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant,
                                              u32 stride, int32_t offset) const {
    check_basis(x);
    check_galois_input(x);
    std::vector<bool> present(entries.size());
    const Ct* first = nullptr;
    for (size_t i = 0; i < entries.size(); ++i) {
        if (!entries[i]) continue;
        check_basis(*entries[i]);
        check_galois_input(*entries[i]);
        present[i] = true;
        if (!first) first = entries[i];
    }
    // check_entries (linalg.rs:545-556): one level, then one exact scale; then x's level
    if (first) {
        for (const Ct* e : entries)
            if (e && e->level() != first->level()) fail(CKH_ERR_LEVEL_MISMATCH, "level mismatch");
        for (const Ct* e : entries)
            if (e && e->scale != first->scale) fail(CKH_ERR_SCALE_MISMATCH, "diagonal entries differ in scale");
        check_level(x.level(), first->level());
    }
    check_bsgs_shape(x.level(), baby, giant, stride, offset, present);
    if (x.level() == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "no levels left");
    relin_key_checked(x.level());
    return present;
}

CtPtr Evaluator::matvec_enc_run(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant, u32 stride,
                                int32_t offset, const std::vector<bool>& present) {
    if (config.lazy_relin)
        return bsgs(
            {{&x, present}}, baby, giant, stride, offset,
            [&](size_t, size_t i, const Ct& xr) { return mul_no_relin(*entries[i], xr); },
            [&](const CtPtr& c) { return relinearize(*c); });
    return bsgs(
        {{&x, present}}, baby, giant, stride, offset,
        [&](size_t, size_t i, const Ct& xr) { return mul(*entries[i], xr); }, [](const CtPtr& c) { return c; });
}

// Evaluator::matvec_enc (linalg.rs:638-665): encrypted 2-component diagonals; lazy_relin: per
// block the tensor products summed at degree 2 and relinearized once, else one mul per product.
CtPtr Evaluator::matvec_enc(const Ct& x, const std::vector<const Ct*>& entries, u32 baby, u32 giant, u32 stride,
                            int32_t offset) {
    const std::vector<bool> present = check_matvec_enc(x, entries, baby, giant, stride, offset);
    return counted([&] { return matvec_enc_run(x, entries, baby, giant, stride, offset, present); });
}

// Device bytes one matvec_enc of this engine holds at its peak, an upper estimate: the shared
// decomposition, the hoisted steps' scratch and rotated copies, a degree-2 product and its sum,
// one key switch's scratch with its output, and the blocks awaiting their giant rotation.
u64 Evaluator::matvec_enc_workspace_bytes(u32 level, u32 baby, u32 giant) const {
    const u64 q = level + 1, k = ctx.np, nd = ctx.ndigits(level), qp = q + k;
    const u64 limbs = nd * qp + q + (baby - 1) * (2 * qp + 2 * k + 5 * q) + 6 * q +
                      (q + nd * qp + 2 * qp + 2 * k + 2 * q) + 2 * q + 4 * q * giant;
    return limbs * ctx.n * 8;
}

// CudaEvaluator::matvec_enc_batch (crates/ckks-train-cuda/src/evaluator.rs:437-519): every job
// checked first (in order), the largest job's working set compared with the device memory
// available now (Device error, nothing done or counted), then the products one after the other
// on the calling thread's stream (one product already fills the device); counts merged only
// when every job succeeded.
std::vector<CtPtr> Evaluator::matvec_enc_batch(const std::vector<const Ct*>& xs,
                                               const std::vector<std::vector<const Ct*>>& entries,
                                               const std::vector<BsgsShape>& shapes) {
    if (xs.size() != entries.size() || xs.size() != shapes.size())
        fail(CKH_ERR_INVALID_PARAMS, "matvec_enc_batch: diagonal sets and vectors differ in number");
    std::vector<std::vector<bool>> present;
    u64 need = 0;
    for (size_t i = 0; i < xs.size(); ++i) {
        const BsgsShape& sh = shapes[i];
        present.push_back(check_matvec_enc(*xs[i], entries[i], sh.baby, sh.giant, sh.stride, sh.offset));
        need = std::max(need, matvec_enc_workspace_bytes(xs[i]->level(), sh.baby, sh.giant));
    }
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const u64 avail = (u64)free_b + ctx.pool_unused_reserve();
    if (need > total_b || need > avail) {
        std::ostringstream os;
        os << "matvec_enc_batch: a product needs " << need / (1 << 20) << " MiB of working memory, more than the "
           << (need > total_b ? "device's total " : "available ") << (need > total_b ? total_b : avail) / (1 << 20)
           << " MiB";
        fail(CKH_ERR_DEVICE, os.str());
    }
    return counted([&] {
        std::vector<CtPtr> out;
        for (size_t i = 0; i < xs.size(); ++i) {
            const BsgsShape& sh = shapes[i];
            out.push_back(matvec_enc_run(*xs[i], entries[i], sh.baby, sh.giant, sh.stride, sh.offset, present[i]));
        }
        return out;
    });
}

// check_mod_raise_input (crates/ckks-train-ckks/src/bootstrap/modraise.rs:94-106).
void Evaluator::check_mod_raise_input(const Ct& ct) const {
    check_basis(ct);
    if (ct.comps.size() != 2) fail(CKH_ERR_INVALID_PARAMS, "mod_raise needs a 2-component ciphertext");
    if (ct.level() != 0) fail(CKH_ERR_LEVEL_MISMATCH, "mod_raise needs a level-0 ciphertext");
}

// Both components raised to max_level, as one program (modraise.rs:131-139).
CtPtr Evaluator::raise_ct(const Ct& ct) {
    const u32 top = ctx.nq - 1;
    const u64 L = top + 1;
    BufPtr out = ctx.alloc(2 * L * ctx.n);
    BufPtr x = ctx.alloc(2 * ctx.n);
    std::vector<RaiseTask> tasks;
    for (u32 i = 0; i < 2; ++i)
        tasks.push_back({ct.comps[i]->data(), top, out->ptr + i * L * ctx.n, x->ptr + i * ctx.n, true});
    ctx.raise_level0(tasks, ctx.stream());
    hip_check(hipGetLastError(), "raise");
    return make_ct({view(out, 0, top, 0, true), view(out, L * ctx.n, top, 0, true)}, ct.scale);
}

// Evaluator::mod_raise_main_key (modraise.rs:77-80).
CtPtr Evaluator::mod_raise_main_key(const Ct& ct) {
    check_mod_raise_input(ct);
    return raise_ct(ct);
}

// Evaluator::mod_raise (modraise.rs:53-69): c_1 switched to the ephemeral key at level 0
// ((c_0 + kb, ka)), both components raised to max_level, c_1 switched back under to_dense;
// sse_switches += 2 on success.
CtPtr Evaluator::mod_raise(const Ct& ct, const Ksk& to_sparse, const Ksk& to_dense) {
    check_mod_raise_input(ct);
    // check_sse_key (modraise.rs:110-128): to_sparse one part at level 0, to_dense dnum parts at
    // max_level, both over this context
    if (to_sparse.ctx != &ctx || to_sparse.max_level != 0 || to_sparse.parts != 1)
        fail(CKH_ERR_INVALID_PARAMS, "to_sparse is not an SSE key of these parameters");
    if (to_dense.ctx != &ctx || to_dense.max_level != ctx.nq - 1 || to_dense.parts != ctx.dnum)
        fail(CKH_ERR_INVALID_PARAMS, "to_dense is not an SSE key of these parameters");
    const u64 n = ctx.n;
    BufPtr s1 = ctx.alloc(2 * n);
    run_ks(ctx, {{ct.comps[1].get(), &to_sparse, ct.comps[0]->data(), nullptr, s1->ptr, s1->ptr + n}});
    const Ct sparse{{view(s1, 0, 0, 0, true), view(s1, n, 0, 0, true)}, ct.scale};
    CtPtr raised = raise_ct(sparse);
    const u32 top = ctx.nq - 1;
    const u64 L = top + 1;
    BufPtr out = ctx.alloc(2 * L * n);
    run_ks(ctx, {{raised->comps[1].get(), &to_dense, raised->comps[0]->data(), nullptr, out->ptr, out->ptr + L * n}});
    CtPtr r = make_ct({view(out, 0, top, 0, true), view(out, L * n, top, 0, true)}, ct.scale);
    count(&ckh_op_counter::sse_switches, 2);
    return r;
}

// HeOps::mul_lazy / relin_lazy (ops.rs:345-363).
CtPtr Evaluator::mul_lazy(const Ct& a, const Ct& b) { return config.lazy_relin ? mul_no_relin(a, b) : mul(a, b); }

CtPtr Evaluator::relin_lazy(const Ct& sum) {
    if (config.lazy_relin) return relinearize(sum);
    return make_ct(sum.comps, sum.scale);
}

// exact_product_scale (evaluator.rs:738-772): the f64 label L with L / q_l == target_scale, the
// nearest of the candidates within 8 ulps of target_scale * q_l (nearest first, above before below).
double Evaluator::exact_product_scale(u32 level, double target_scale) const {
    if (level > ctx.nq - 1) fail(CKH_ERR_INVALID_PARAMS, "level exceeds max_level");
    if (!(std::isfinite(target_scale) && target_scale > 0.0))
        fail(CKH_ERR_INVALID_PARAMS, "target scale must be positive and finite");
    if (level == 0) fail(CKH_ERR_NO_LEVELS_LEFT, "no levels left");
    const double q = (double)ctx.primes[level];
    const double l0 = target_scale * q;
    if (!std::isfinite(l0)) fail(CKH_ERR_INVALID_PARAMS, "target scale times q_l overflows");
    u64 bits;
    std::memcpy(&bits, &l0, 8);
    for (u64 k = 0; k <= 8; ++k)
        for (u64 cand : {bits + k, bits - k}) {
            double l;
            std::memcpy(&l, &cand, 8);
            if (l / q == target_scale) return l;
        }
    fail(CKH_ERR_SCALE_MISMATCH, "no label within 8 ulps divides to the target scale");
}

// check_plain_scale (evaluator.rs:108-119).
static void check_plain_scale(double sp) {
    if (sp < 1048576.0) fail(CKH_ERR_INVALID_PARAMS, "plaintext scale below 2^20");
}

// Evaluator::target_plain_scale (evaluator.rs:341-345).
double Evaluator::target_plain_scale(const Ct& ct, double target_scale) const {
    const double sp = exact_product_scale(ct.level(), target_scale) / ct.scale;
    check_plain_scale(sp);
    return sp;
}

// Evaluator::mul_scalar (evaluator.rs:369-383) with scalar_product_parts (775-797): every
// component times the integer k = round(value * L / scale) (negated for k < 0), labelled L.
CtPtr Evaluator::mul_scalar(const Ct& ct, double value, double target_scale) {
    check_basis(ct);
    if (!std::isfinite(value)) fail(CKH_ERR_INVALID_PARAMS, "scalar must be finite");
    const double label = exact_product_scale(ct.level(), target_scale);
    check_plain_scale(label / ct.scale);
    const double k = std::round(value * (label / ct.scale));
    if (std::fabs(k) >= 4611686018427387904.0) fail(CKH_ERR_INVALID_PARAMS, "scalar integer at or above 2^62");
    std::vector<PolyPtr> out;
    for (auto& p : ct.comps) out.push_back(poly_scalar_mul(ctx, *p, (u64)std::fabs(k), k < 0.0));
    CtPtr r = make_ct(std::move(out), label);
    count(&ckh_op_counter::ct_pt_mults);
    return r;
}

// Evaluator::mul_plain_to_scale (evaluator.rs:399-418).
CtPtr Evaluator::mul_plain_to_scale(const Ct& a, const Poly& pt, double pt_scale, double target_scale) {
    check_basis(a);
    check_plain(ctx, pt);
    check_level(a.level(), pt.level);
    const double label = exact_product_scale(a.level(), target_scale);
    check_plain_scale(label / a.scale);
    if (pt_scale != label / a.scale) fail(CKH_ERR_SCALE_MISMATCH, "plaintext scale is not the target plaintext scale");
    std::vector<PolyPtr> out;
    for (auto& c : a.comps) out.push_back(poly_pointwise(ctx, PW_MUL, *c, &pt));
    CtPtr r = make_ct(std::move(out), label);
    count(&ckh_op_counter::ct_pt_mults);
    return r;
}

// ------------------------------------------------------------------ CoeffToSlot / SlotToCoeff

// Evaluator::check_stages (linear.rs:861-898), before any work: stage k at level - k, every
// diagonal scale exactly the one that takes the planned input scale (scale, then each stage's
// scale) to the stage's pre-rescale label, a second diagonal set of the same layout and scale,
// and every rotation key (the shape checks of the BSGS product are made here too, so that no
// stage fails after an earlier one ran).
void Evaluator::check_stages(const std::vector<LinearStage>& stages, u32 level, double scale) const {
    if (stages.empty()) fail(CKH_ERR_INVALID_PARAMS, "no stages");
    double s_in = scale;
    for (size_t k = 0; k < stages.size(); ++k) {
        const LinearStage& lt = stages[k];
        if ((u64)lt.level + k != level) fail(CKH_ERR_LEVEL_MISMATCH, "stage levels must descend by one from the input's");
        const double label = exact_product_scale(lt.level, lt.scale);
        bool any = false;
        for (const Poly* w : lt.diags) any = any || w;
        if (!any) fail(CKH_ERR_INVALID_PARAMS, "a stage without diagonals");
        if (lt.pt_scale != label / s_in) fail(CKH_ERR_SCALE_MISMATCH, "diagonal scale does not give the stage scale");
        const BsgsShape& a = lt.shape;
        if (lt.has_second) {
            const BsgsShape& b = lt.second_shape;
            if (b.baby != a.baby || b.giant != a.giant || b.stride != a.stride || b.offset != a.offset)
                fail(CKH_ERR_INVALID_PARAMS, "second diagonals of a different layout");
            bool any2 = false;
            for (const Poly* w : lt.second) any2 = any2 || w;
            if (!any2 || lt.second_pt_scale != lt.pt_scale)
                fail(CKH_ERR_SCALE_MISMATCH, "second diagonals of a different scale");
        }
        const std::vector<bool> all((size_t)a.baby * a.giant, true);
        check_bsgs_shape(lt.level, a.baby, a.giant, a.stride, a.offset, all);
        if (lt.diags.size() != all.size()) fail(CKH_ERR_INVALID_PARAMS, "entries must hold baby * giant diagonals");
        if (lt.has_second && lt.second.size() != all.size())
            fail(CKH_ERR_INVALID_PARAMS, "entries must hold baby * giant diagonals");
        for (const auto* set : {&lt.diags, &lt.second})
            for (const Poly* w : *set) {
                if (!w) continue;
                check_plain(ctx, *w);
                check_level(lt.level, w->level);
            }
        s_in = lt.scale;
    }
}

// Evaluator::apply_stage (linear.rs:900-919): the BSGS sum of diags * x (and second * x2),
// labelled with the exact pre-rescale scale (within 1e-15 of the f64 product of the scales) and
// rescaled to the stage's scale.
CtPtr Evaluator::apply_stage(const LinearStage& lt, const Ct& x, const Ct* x2) {
    std::vector<const Ct*> xs{&x};
    std::vector<std::vector<const Poly*>> entries{lt.diags};
    if (lt.has_second && x2) {
        xs.push_back(x2);
        entries.push_back(lt.second);
    }
    const BsgsShape& a = lt.shape;
    CtPtr y = matvec_plain_sum(xs, entries, lt.pt_scale, a.baby, a.giant, a.stride, a.offset);
    const double label = exact_product_scale(lt.level, lt.scale);
    if (std::fabs((y->scale - label) / label) > 1e-15) fail(CKH_ERR_SCALE_MISMATCH, "stage product scale is not its label");
    return rescale(*relabel(*y, label));
}

// Evaluator::mul_by_x_half (linear.rs:921-945): every component times X^(N/2) (NTT form,
// pointwise), which multiplies slot j by i^(3^j). No key, no level, not counted.
CtPtr Evaluator::mul_by_x_half(const Ct& ct) {
    check_basis(ct);
    std::call_once(mono_once_, [&] {
        std::vector<u64> host((u64)ctx.nq * ctx.n, 0);
        for (u32 i = 0; i < ctx.nq; ++i) host[(u64)i * ctx.n + ctx.n / 2] = 1;
        PolyPtr coeff = poly_upload(ctx, host.data(), ctx.nq - 1, 0, false);
        PolyPtr m = poly_transform(ctx, *coeff, true);
        hip_check(hipStreamSynchronize(ctx.stream()), "monomial");  // used from any stream afterwards
        mono_ = m;
    });
    std::vector<PolyPtr> out;
    for (auto& p : ct.comps) {
        PolyPtr mono = view(mono_->buf, mono_->off, p->level, 0, true);
        out.push_back(poly_pointwise(ctx, PW_MUL, *p, mono.get()));
    }
    return make_ct(std::move(out), ct.scale);
}

// Evaluator::coeff_to_slot (linear.rs:776-806): the stages in order (the last also applied to
// the conjugate of its input), then re = x + conj(x), im = (conj(x) - x) X^(N/2).
std::pair<CtPtr, CtPtr> Evaluator::coeff_to_slot(const Ct& ct, const std::vector<LinearStage>& stages) {
    check_basis(ct);
    check_galois_input(ct);
    if (stages.empty()) fail(CKH_ERR_INVALID_PARAMS, "no stages");
    for (auto& lt : stages)
        if (lt.slot_to_coeff) fail(CKH_ERR_INVALID_PARAMS, "coeff_to_slot: a stage of the SlotToCoeff direction");
    bool order_ok = stages.back().has_second;
    for (size_t k = 0; k + 1 < stages.size(); ++k) order_ok = order_ok && !stages[k].has_second;
    if (!order_ok) fail(CKH_ERR_INVALID_PARAMS, "coeff_to_slot: only the last stage (and it must) takes the conjugate");
    check_stages(stages, ct.level(), ct.scale);
    galois_key(2 * ctx.n - 1, stages.back().level);  // the conjugation key, at the highest level it is used
    return counted([&] {
        CtPtr x = make_ct(ct.comps, ct.scale);
        for (auto& lt : stages) {
            CtPtr xc = lt.has_second ? conjugate(*x) : nullptr;
            x = apply_stage(lt, *x, xc.get());
        }
        CtPtr xc = conjugate(*x);
        CtPtr re = add(*x, *xc, false);
        CtPtr im = mul_by_x_half(*add(*xc, *x, true));
        return std::make_pair(re, im);
    });
}

// Evaluator::slot_to_coeff (linear.rs:824-859): the first stage takes re and im, the others the
// previous stage's output.
CtPtr Evaluator::slot_to_coeff(const Ct& re, const Ct& im, const std::vector<LinearStage>& stages) {
    check_basis(re);
    check_galois_input(re);
    check_basis(im);
    check_galois_input(im);
    for (auto& lt : stages)
        if (!lt.slot_to_coeff) fail(CKH_ERR_INVALID_PARAMS, "slot_to_coeff: a stage of the CoeffToSlot direction");
    bool order_ok = !stages.empty() && stages[0].has_second;
    for (size_t k = 1; k < stages.size(); ++k) order_ok = order_ok && !stages[k].has_second;
    if (!order_ok) fail(CKH_ERR_INVALID_PARAMS, "slot_to_coeff: only the first stage (and it must) takes two inputs");
    if (re.level() != im.level()) fail(CKH_ERR_LEVEL_MISMATCH, "re and im at different levels");
    if (re.scale != im.scale) fail(CKH_ERR_SCALE_MISMATCH, "re and im of different scales");
    check_stages(stages, re.level(), re.scale);
    return counted([&] {
        CtPtr x = apply_stage(stages[0], re, &im);
        for (size_t k = 1; k < stages.size(); ++k) x = apply_stage(stages[k], *x, nullptr);
        return x;
    });
}

}  // namespace ckh
