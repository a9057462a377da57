// Host runtime: context construction (device tables), stream-ordered memory, descriptor
// upload, and the batched key-switching engine.
#include "runtime.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <sstream>
#include <unordered_map>

#include "hostmath.h"
#include "conv_tables.h"
#include "ntt_tables.h"

namespace ckh {

[[noreturn]] void fail(ckh_status st, const std::string& msg) { throw Error{st, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::ostringstream os;
        os << what << ": " << hipGetErrorString(e);
        fail(CKH_ERR_DEVICE, os.str());
    }
}

Buffer::~Buffer() {
    if (ctx) ctx->free_buffer(this);
    // `keep` is released after the free has been queued on the context's stream.
}

u64 Ksk::poly_words() const { return (u64)(max_level + 1 + ctx->np) * ctx->n; }

// ------------------------------------------------------------------ context

namespace {
thread_local std::unordered_map<const Context*, u32> tl_stream;
}

Context::Context(int ordinal, const ckh_params& p, u32 n_streams) {
    if (p.log_n < 1 || p.log_n > 17) fail(CKH_ERR_INVALID_PARAMS, "log_n must be in 1..=17");
    if (p.nq == 0 || p.np == 0 || !p.q || !p.p) fail(CKH_ERR_INVALID_PARAMS, "q and p must be non-empty");
    if (p.dnum == 0 || p.nq % p.dnum != 0)
        fail(CKH_ERR_INVALID_PARAMS, "dnum must divide the number of q primes");
    log_n = p.log_n;
    n = 1ull << log_n;
    nq = p.nq;
    np = p.np;
    dnum = p.dnum;
    alpha = nq / dnum;
    if (alpha > MAX_SRC || np > MAX_SRC)
        fail(CKH_ERR_INVALID_PARAMS, "at most 16 primes per digit and 16 special primes");
    primes.assign(p.q, p.q + nq);
    primes.insert(primes.end(), p.p, p.p + np);
    for (size_t i = 0; i < primes.size(); ++i) {
        u64 q = primes[i];
        if (q < 2 || q >= (1ull << 60) || (q - 1) % (2 * n) != 0 || !hm::is_prime(q))
            fail(CKH_ERR_INVALID_PARAMS, "every modulus must be a prime < 2^60 with q = 1 mod 2N");
        for (size_t j = 0; j < i; ++j)
            if (primes[j] == q) fail(CKH_ERR_INVALID_PARAMS, "duplicate modulus");
    }
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (ordinal < 0 || ordinal >= count) fail(CKH_ERR_DEVICE, "no such device ordinal");
    device = ordinal;
    hip_check(hipSetDevice(device), "hipSetDevice");
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
        uint64_t thr = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    if (n_streams == 0) n_streams = 1;
    n_user_streams_ = n_streams;
    for (u32 i = 0; i < n_streams + SERVICE_STREAMS; ++i) {
        hipStream_t s;
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        streams_.push_back(s);
    }
    hip_check(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking), "hipStreamCreate");
    // The pinned staging ring of h2d and a stock of its events, made here so that uploads
    // (descriptors of the no-allocation key switch included) allocate nothing later.
    {
        ring_cap_ = (size_t)16 << 20;
        void* rp = nullptr;
        hip_check(hipHostMalloc(&rp, ring_cap_, hipHostMallocDefault), "hipHostMalloc(staging ring)");
        ring_ = (uint8_t*)rp;
        for (int i = 0; i < 32; ++i) {
            hipEvent_t ev;
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
            ring_events_.push_back(ev);
        }
    }
    build_tables();
    static_scratch = static_scratch_address(device);
    {
        void* e = nullptr;
        hip_check(hipMalloc(&e, 256), "hipMalloc(error word)");
        perm_allocs_.push_back(e);
        hip_check(hipMemset(e, 0, 256), "hipMemset");
        d_err = (u32*)e;
        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "attribute");
        fused.nx = xcc_layout_check(streams_[0]);
        cu_count = (u32)cus;
        fused.grid = cu_count * std::max<u32>(1, fused_occupancy());
        fused.err = d_err;
    }
}

Context::~Context() {
    (void)hipSetDevice(device);
    for (auto s : streams_) (void)hipStreamSynchronize(s);
    for (auto& b : ring_busy_) (void)hipEventDestroy(b.ev);
    for (auto e : ring_events_) (void)hipEventDestroy(e);
    if (ring_) (void)hipHostFree(ring_);
    for (auto& pl : plans_) {
        if (pl->done) (void)hipEventDestroy(pl->done);
        if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
        if (pl->desc) (void)hipFreeAsync(pl->desc, pl->s);
    }
    plans_.clear();
    for (auto& pl : retired_) {  // every stream was synchronized above
        (void)hipEventDestroy(pl->done);
        if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    }
    retired_.clear();
    for (auto s : streams_) (void)hipStreamSynchronize(s);
    cache_release(0);
    for (hipEvent_t e : cache_events_) (void)hipEventDestroy(e);
    for (void* a : perm_allocs_) (void)hipFree(a);
    for (auto s : streams_) (void)hipStreamDestroy(s);
    if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
}

// A thread's choice is kept per context address; an entry left by a destroyed context at the
// same address is out of range or harmless.
hipStream_t Context::stream() const {
    auto it = tl_stream.find(this);
    const u32 i = it == tl_stream.end() ? 0 : it->second;
    return streams_[i < streams_.size() ? i : 0];
}

void Context::set_stream(u32 index) {
    if (index >= n_user_streams_) fail(CKH_ERR_INVALID_PARAMS, "stream index out of range");
    tl_stream[this] = index;
}

void Context::set_service_stream(u32 k) {
    tl_stream[this] = n_user_streams_ + k % std::min(std::max(service_streams, 1u), SERVICE_STREAMS);
}

// Block size classes: 4 KiB up to 4 KiB, then four per power of two (at most 25 % slack).
static size_t size_class(size_t b) {
    if (b <= 4096) return 4096;
    const int k = 63 - __builtin_clzll((unsigned long long)(b - 1));  // 2^k < b <= 2^(k+1)
    const size_t base = (size_t)1 << k, step = base / 4;
    const size_t c = base + ((b - base + step - 1) / step) * step;
    return (c + 4095) & ~(size_t)4095;
}

// A cached block of class cls for stream s: one last used on s (stream order makes it free for
// the work queued after), else one whose event has completed; null if none.
void* Context::cache_take(size_t& cls, hipStream_t s) {
    std::lock_guard<std::mutex> lk(cache_mu_);
    // The class itself first, then larger classes up to twice its size (a fitting free block is
    // reused instead of growing the cache with another class); within a class, a block last used
    // on s, else one whose event has completed.
    for (auto it = cache_.lower_bound(cls); it != cache_.end() && it->first <= 2 * cls; ++it) {
        std::vector<Cached>& v = it->second;
        size_t pick = v.size();
        for (size_t i = 0; i < v.size() && pick == v.size(); ++i)
            if (v[i].s == s) pick = i;
        for (size_t i = 0; i < v.size() && pick == v.size(); ++i)
            if (hipEventQuery(v[i].ready) == hipSuccess) pick = i;
        if (pick == v.size()) continue;
        void* p = v[pick].p;
        cache_events_.push_back(v[pick].ready);
        v[pick] = v.back();
        v.pop_back();
        cls = it->first;
        cache_bytes_ -= cls;
        return p;
    }
    return nullptr;
}

// Caches a block whose last use is queued on s (over the release threshold it goes back to the
// driver once that use has finished).
void Context::cache_put(void* p, size_t cls, hipStream_t s) {
    std::lock_guard<std::mutex> lk(cache_mu_);
    hipEvent_t ev = nullptr;
    if (!cache_events_.empty()) {
        ev = cache_events_.back();
        cache_events_.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        ev = nullptr;
    }
    if (cache_bytes_ + cls > cache_limit_ || !ev || hipEventRecord(ev, s) != hipSuccess) {
        (void)hipGetLastError();
        if (ev) cache_events_.push_back(ev);
        (void)hipStreamSynchronize(s);
        (void)hipFree(p);
        return;
    }
    cache_[cls].push_back({p, s, ev});
    cache_bytes_ += cls;
}

// Returns cached blocks to the driver, largest first, until at most keep_bytes stay cached.
void Context::cache_release(u64 keep_bytes) {
    std::lock_guard<std::mutex> lk(cache_mu_);
    for (auto it = cache_.rbegin(); it != cache_.rend() && cache_bytes_ > keep_bytes; ++it) {
        std::vector<Cached>& v = it->second;
        while (!v.empty() && cache_bytes_ > keep_bytes) {
            (void)hipEventSynchronize(v.back().ready);
            (void)hipFree(v.back().p);
            cache_events_.push_back(v.back().ready);
            v.pop_back();
            cache_bytes_ -= it->first;
        }
    }
}

u64 Context::pool_used_memory() {
    if (pool_mode == 1) return live_bytes.load();
    hipMemPool_t mp;
    hip_check(hipDeviceGetDefaultMemPool(&mp, device), "hipDeviceGetDefaultMemPool");
    uint64_t v = 0;
    hip_check(hipMemPoolGetAttribute(mp, hipMemPoolAttrUsedMemCurrent, &v), "hipMemPoolGetAttribute");
    return v;
}

u64 Context::pool_unused_reserve() {
    if (pool_mode == 1) {
        std::lock_guard<std::mutex> lk(cache_mu_);
        return cache_bytes_;
    }
    hipMemPool_t mp;
    hip_check(hipDeviceGetDefaultMemPool(&mp, device), "hipDeviceGetDefaultMemPool");
    uint64_t used = 0, reserved = 0;
    hip_check(hipMemPoolGetAttribute(mp, hipMemPoolAttrUsedMemCurrent, &used), "hipMemPoolGetAttribute");
    hip_check(hipMemPoolGetAttribute(mp, hipMemPoolAttrReservedMemCurrent, &reserved), "hipMemPoolGetAttribute");
    return reserved > used ? reserved - used : 0;
}

void Context::set_pool_release_threshold(u64 bytes) {
    {
        std::lock_guard<std::mutex> lk(cache_mu_);
        cache_limit_ = bytes;
    }
    uint64_t v = bytes;
    hipMemPool_t dp;
    hip_check(hipDeviceGetDefaultMemPool(&dp, device), "hipDeviceGetDefaultMemPool");
    hip_check(hipMemPoolSetAttribute(dp, hipMemPoolAttrReleaseThreshold, &v), "pool threshold");
}

void Context::synchronize_all() {
    {
        std::lock_guard<std::mutex> lk(plan_mu_);
        for (auto st : streams_) hip_check(hipStreamSynchronize(st), "synchronize");
    }
    check_device_error(stream());
}

void Context::check_device_error(hipStream_t s) {
    if (!d_err) return;
    u32 v = 0;
    hip_check(hipMemcpyAsync(&v, d_err, sizeof(u32), hipMemcpyDeviceToHost, s), "error word");
    hip_check(hipStreamSynchronize(s), "error word");
    if (v) {
        hip_check(hipMemsetAsync(d_err, 0, sizeof(u32), s), "error word");
        hip_check(hipStreamSynchronize(s), "error word");
        fail(CKH_ERR_DEVICE, "a one-launch NTT waited past its bound; results of work since the last check are unreliable");
    }
}

void Context::trim_pool(u64 keep_bytes) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    synchronize_all();
    cache_release(keep_bytes);
    hipMemPool_t dp;
    hip_check(hipDeviceGetDefaultMemPool(&dp, device), "hipDeviceGetDefaultMemPool");
    hip_check(hipMemPoolTrimTo(dp, keep_bytes), "hipMemPoolTrimTo");
}

BufPtr Context::alloc(size_t words, hipStream_t s) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(words, 1) * sizeof(u64), cls = 0;
    if (fail_alloc_after.load() >= 0 && fail_alloc_after.fetch_sub(1) == 0)
        fail(CKH_ERR_DEVICE, "out of device memory (injected allocation failure)");
    hipError_t e;
    if (pool_mode == 1) {
        cls = size_class(bytes);
        p = cache_take(cls, s);
        e = hipSuccess;
        if (!p) {
            const auto t0 = std::chrono::steady_clock::now();
            e = hipMalloc(&p, cls);
            ++stat_mallocs;
            stat_malloc_us += (u64)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
            if (e != hipSuccess) {  // cached blocks back to the driver, then once more
                (void)hipGetLastError();
                ++stat_oom;
                {
                    std::lock_guard<std::mutex> lk(plan_mu_);  // as in synchronize_all
                    for (auto st : streams_) (void)hipStreamSynchronize(st);
                }
                // enough for this block and as much again (largest classes first), so that a
                // device near full does not empty and regrow the whole cache on every miss
                u64 cached;
                {
                    std::lock_guard<std::mutex> lk(cache_mu_);
                    cached = cache_bytes_;
                }
                cache_release(cached > 2 * cls ? cached - 2 * cls : 0);
                e = hipMalloc(&p, cls);
                if (e != hipSuccess) {
                    (void)hipGetLastError();
                    cache_release(0);
                    e = hipMalloc(&p, cls);
                }
            }
        }
        bytes = cls;
    } else {
        e = hipMallocAsync(&p, bytes, s);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::ostringstream os;
        os << "out of device memory allocating " << bytes << " bytes";
        fail(CKH_ERR_DEVICE, os.str());
    }
    live_bytes += bytes;
    auto b = std::make_shared<Buffer>();
    b->ctx = this;
    b->keep = shared_from_this();
    b->ptr = (u64*)p;
    b->words = words;
    b->stream = s;
    b->cls = cls;
    return b;
}

void Context::free_buffer(Buffer* b) {
    if (!b->ptr) return;
    (void)hipSetDevice(device);
    // Streams whose queued work may still use the buffer: the releasing thread's and every other
    // user recorded at use; the free (on the allocation stream) is ordered after each of them.
    std::vector<hipStream_t> order = free_order(b->stream, stream(), b->users);
    if (!order.empty()) {
        // Holding the plan lock keeps these records off a graph capture in progress.
        std::lock_guard<std::mutex> lk(plan_mu_);
        for (hipStream_t r : order) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
                if (hipEventRecord(ev, r) == hipSuccess) (void)hipStreamWaitEvent(b->stream, ev, 0);
                (void)hipEventDestroy(ev);
            } else {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(r);
            }
        }
    }
    if (b->cls)
        cache_put(b->ptr, b->cls, b->stream);
    else
        (void)hipFreeAsync(b->ptr, b->stream);
    live_bytes -= b->cls ? b->cls : std::max<size_t>(b->words, 1) * sizeof(u64);
    b->ptr = nullptr;
}

BufPtr Context::upload_blob(const void* host, size_t bytes, hipStream_t s) {
    BufPtr b = alloc((bytes + 7) / 8, s);
    h2d(b->ptr, host, bytes, s);
    return b;
}

// A ring slot of len bytes: slots are handed out in address order, wrapping at the end; a slot
// still read by an unfinished copy (the oldest busy slots, from the previous lap) is waited for.
size_t Context::ring_reserve(size_t len) {
    len = (len + 255) & ~(size_t)255;
    if (ring_head_ + len > ring_cap_) ring_head_ = 0;
    const size_t off = ring_head_;
    ring_head_ += len;
    while (!ring_busy_.empty()) {
        Staged& b = ring_busy_.front();
        const bool overlaps = b.off < off + len && off < b.off + b.len;
        if (!overlaps && hipEventQuery(b.ev) != hipSuccess) break;
        if (overlaps) hip_check(hipEventSynchronize(b.ev), "staging ring");
        ring_events_.push_back(b.ev);
        ring_busy_.pop_front();
    }
    for (const Staged& b : ring_busy_)  // a later busy slot can overlap only after a wrap
        if (b.off < off + len && off < b.off + b.len) hip_check(hipEventSynchronize(b.ev), "staging ring");
    return off;
}

void Context::h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return;
    std::lock_guard<std::mutex> lk(ring_mu_);
    const size_t chunk_max = ring_cap_ / 4;
    for (size_t done = 0; done < bytes;) {
        const size_t len = std::min(chunk_max, bytes - done);
        const size_t off = ring_reserve(len);
        std::memcpy(ring_ + off, (const uint8_t*)src + done, len);
        hip_check(hipMemcpyAsync((uint8_t*)dst + done, ring_ + off, len, hipMemcpyHostToDevice, s), "upload");
        hipEvent_t ev;
        if (ring_events_.empty()) {
            hip_check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
        } else {
            ev = ring_events_.back();
            ring_events_.pop_back();
        }
        hip_check(hipEventRecord(ev, s), "event");
        ring_busy_.push_back({off, len, ev});
        done += len;
    }
}

template <class T>
T* Context::dev_copy(const std::vector<T>& v) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    hip_check(hipMalloc(&p, bytes), "hipMalloc(tables)");
    perm_allocs_.push_back(p);
    if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "table upload");
    return (T*)p;
}

std::vector<std::pair<u32, u32>> Context::digit_ranges(u32 level) const {
    std::vector<std::pair<u32, u32>> r;
    for (u32 j = 0; j < dnum; ++j) {
        u32 s = j * alpha, e = std::min((j + 1) * alpha, level + 1);
        if (s > level) break;
        r.push_back({s, e});
    }
    return r;
}

u32 Context::ndigits(u32 level) const { return (u32)digit_ranges(level).size(); }

// Device tables: per-prime constants and twiddles (NttTable::new, ntt.rs:89-122), and the
// conversion plans of ModUp (per level and digit), ModDown (per level) and rescale (per level).
void Context::build_tables() {
    HostNttTables t = build_ntt_tables(primes, log_n, ntt_fast_supported(log_n));
    d_pc = dev_copy(t.pc);
    ntt.log_n = log_n;
    ntt.pc = d_pc;
    ntt.fwd = dev_copy(t.fwd);
    ntt.inv = dev_copy(t.inv);
    ntt.fwd_f = dev_copy(t.fwd_f);
    ntt.inv_f = dev_copy(t.inv_f);
    if (!t.rf.empty()) {
        ntt.rf = dev_copy(t.rf);
        ntt.ri = dev_copy(t.ri);
    }

    HostConvTables ct;
    auto& plans = ct.plans;
    auto add_plan = [&](const std::vector<u32>& src, const std::vector<std::pair<u32, u32>>& tg, u32 limb0,
                        bool share_tgt_first, u32 shared_first) {
        return ct.add_plan(primes, src, tg, limb0, share_tgt_first, shared_first);
    };

    modup_plan_.assign(nq, {});
    modup_maxtgt_.assign(nq, 0);
    for (u32 l = 0; l < nq; ++l) {
        for (auto [s, e] : digit_ranges(l)) {
            std::vector<u32> src;
            for (u32 i = s; i < e; ++i) src.push_back(i);
            std::vector<std::pair<u32, u32>> tg;
            for (u32 i = 0; i <= l; ++i)
                if (i < s || i >= e) tg.push_back({i, i});
            for (u32 t = 0; t < np; ++t) tg.push_back({nq + t, l + 1 + t});
            modup_maxtgt_[l] = std::max(modup_maxtgt_[l], (u32)tg.size());
            modup_plan_[l].push_back(add_plan(src, tg, s, false, 0));
        }
    }
    // ModDown: P -> Q_l; the targets of level l are the first l + 1 of one shared list.
    {
        std::vector<u32> src;
        for (u32 t = 0; t < np; ++t) src.push_back(nq + t);
        std::vector<std::pair<u32, u32>> all;
        for (u32 i = 0; i < nq; ++i) all.push_back({i, i});
        u32 top = add_plan(src, all, 0, false, 0);
        u32 first = plans[top].tgt_first;
        moddown_plan_.assign(nq, 0);
        for (u32 l = 0; l < nq; ++l) {
            std::vector<std::pair<u32, u32>> tg(all.begin(), all.begin() + l + 1);
            moddown_plan_[l] = add_plan(src, tg, 0, true, first);
        }
        pinv_.resize(nq);
        pinv_s_.resize(nq);
        for (u32 i = 0; i < nq; ++i) {
            u64 pm = hm::prod_mod(primes.data() + nq, np, primes[i]);
            pinv_[i] = hm::inv(pm, primes[i]);
            pinv_s_[i] = hm::shoup(pinv_[i], primes[i]);
        }
    }
    // Rescale: q_l -> q_0..q_{l-1}.
    rescale_plan_.assign(nq, 0);
    qlinv_.assign(nq, {});
    qlinv_s_.assign(nq, {});
    for (u32 l = 1; l < nq; ++l) {
        std::vector<std::pair<u32, u32>> tg;
        for (u32 i = 0; i < l; ++i) tg.push_back({i, i});
        rescale_plan_[l] = add_plan({l}, tg, 0, false, 0);
        for (u32 i = 0; i < l; ++i) {
            u64 v = hm::inv(primes[l] % primes[i], primes[i]);
            qlinv_[l].push_back(v);
            qlinv_s_[l].push_back(hm::shoup(v, primes[i]));
        }
    }
    // Raise from level 0: q_0 -> q_1..q_top (the targets of top are the first top of one list;
    // a single source, so the centered lift needs no overflow estimate).
    raise_plan_.assign(nq, 0);
    if (nq > 1) {
        std::vector<std::pair<u32, u32>> all;
        for (u32 i = 1; i < nq; ++i) all.push_back({i, i});
        const u32 full = add_plan({0}, all, 0, false, 0);
        const u32 first = plans[full].tgt_first;
        for (u32 l = 1; l < nq; ++l) {
            std::vector<std::pair<u32, u32>> tg(all.begin(), all.begin() + l);
            raise_plan_[l] = add_plan({0}, tg, 0, true, first);
        }
    }
    conv.src = dev_copy(ct.srcs);
    conv.yf = dev_copy(ct.yform(t.pc));
    conv_src_first_.clear();
    for (const ConvPlan& pl : plans) conv_src_first_.push_back(pl.src_first);
    conv.tgt = dev_copy(ct.tgts);
    conv.plans = dev_copy(plans);
    conv.pc = d_pc;
    conv.log_n = log_n;
    build_mf_tables(ct);
    hip_check(hipDeviceSynchronize(), "table upload");
}

// The matrix-core form of every conversion plan (HostConvTables::build_mf), on the device.
void Context::build_mf_tables(HostConvTables& ct) {
    conv.mf_ok = ct.build_mf(primes, n) ? 1 : 0;
    conv.mf.plans = dev_copy(ct.mf_plans);
    conv.mf.a = (const mf_v4i*)dev_copy(ct.mf_a);
    conv.mf.tgt = dev_copy(ct.mf_tgt);
}

// ------------------------------------------------------------------ descriptor blobs

namespace {
constexpr u32 MAX_ROWS = 65535;
}  // namespace

// The inverse NTT of refs whose aux names a conversion source: two-pass register-radix kernels,
// the outputs in that source's y form.
static void launch_ntt_y(const Context& c, const LimbRef* d_refs, u32 nrefs, hipStream_t s) {
    u32 chunk = c.ntt_chunk_inv;
    chunk = (chunk == 0 || chunk > MAX_ROWS) ? MAX_ROWS : chunk;
    for (u32 i = 0; i < nrefs; i += chunk) ntt_fast_inverse(c.ntt, d_refs + i, std::min(chunk, nrefs - i), s, c.conv.yf);
}

void launch_ntt(const Context& c, const LimbRef* d_refs, u32 nrefs, bool forward, hipStream_t s) {
    u32 chunk = forward ? c.ntt_chunk_fwd : c.ntt_chunk_inv;
    chunk = (chunk == 0 || chunk > MAX_ROWS) ? MAX_ROWS : chunk;
    for (u32 i = 0; i < nrefs; i += chunk) {
        u32 cnt = std::min(chunk, nrefs - i);
        if (forward)
            ntt_forward(c.ntt, d_refs + i, cnt, s);
        else
            ntt_inverse(c.ntt, d_refs + i, cnt, s);
    }
}

static void launch_copy(const LimbRef* d, u32 nrefs, u32 log_n, hipStream_t s) {
    for (u32 i = 0; i < nrefs; i += MAX_ROWS) copy_launch(d + i, std::min(MAX_ROWS, nrefs - i), log_n, s);
}

static void launch_conv(const ConvTables& t, const ConvJob* d, u32 njobs, u32 max_tgt, u32 max_src, u32 variant,
                        u32 tg, bool yform, hipStream_t s) {
    for (u32 i = 0; i < njobs; i += MAX_ROWS)
        conv_launch(t, d + i, std::min(MAX_ROWS, njobs - i), max_tgt, max_src, variant, tg, yform, s);
}

static void launch_ip(const IpJob* d, u32 njobs, const PrimeConst* pc, u32 log_n, hipStream_t s) {
    for (u32 i = 0; i < njobs; i += MAX_ROWS) inner_product_launch(d + i, std::min(MAX_ROWS, njobs - i), pc, log_n, s);
}

static void launch_finish(const FinishJob* d, u32 njobs, const PrimeConst* pc, u32 log_n, hipStream_t s) {
    for (u32 i = 0; i < njobs; i += MAX_ROWS) finish_launch(d + i, std::min(MAX_ROWS, njobs - i), pc, log_n, s);
}

// ------------------------------------------------------------------ engine
//
// Each stage builds its descriptors into a Program (host bytes + steps); execute() copies the
// bytes to the device in one transfer and launches the steps. A Program's steps depend only on
// the shapes of its tasks, so a captured launch sequence can be replayed with new bytes.

void Program::step(StepKind k, size_t off, u32 count, size_t off2, u32 a, u32 b) {
    if (count) steps.push_back({k, count, off, off2, a, b});
}

void Program::step_ntt(StepKind k, size_t off, u32 count, size_t off2, bool one_launch) {
    if (!count) return;
    const size_t ctl = one_launch ? add(std::vector<u32>(fused_ctl_words(count), 0u)) : 0;
    steps.push_back({k, count, off, off2, 0, 0});
    steps.back().ctl = ctl;
}

void Context::run(const Program& p, const uint8_t* d, hipStream_t s) const {
    for (const Program::Step& st : p.steps) {
        const void* x = d + st.off;
        switch (st.kind) {
            case S_NTT_FWD:
                if (st.ctl && fused_usable())
                    ntt_fast_forward_fused(ntt, (const LimbRef*)x, st.count, nullptr, (u32*)(d + st.ctl), fused, s);
                else
                    launch_ntt(*this, (const LimbRef*)x, st.count, true, s);
                break;
            case S_NTT_INV:
                if (st.ctl && fused_usable())
                    ntt_fast_inverse_fused(ntt, (const LimbRef*)x, st.count, (u32*)(d + st.ctl), fused, s);
                else
                    launch_ntt(*this, (const LimbRef*)x, st.count, false, s);
                break;
            case S_NTT_INV_Y: launch_ntt_y(*this, (const LimbRef*)x, st.count, s); break;
            case S_CONV:
                launch_conv(conv, (const ConvJob*)x, st.count, st.a, st.b, st.sched ? (st.sched >> 1) & 127 : conv_variant,
                            st.sched ? st.sched >> 8 : conv_tg, st.yform, s);
                break;
            case S_COPY: launch_copy((const LimbRef*)x, st.count, log_n, s); break;
            case S_IP: launch_ip((const IpJob*)x, st.count, d_pc, log_n, s); break;
            case S_NTT_FWD_FIN:
                if (st.ctl && fused_usable())
                    ntt_fast_forward_fused(ntt, (const LimbRef*)x, st.count, (const FinishJob*)(d + st.off2),
                                           (u32*)(d + st.ctl), fused, s);
                else
                    launch_ntt_finish((const LimbRef*)x, (const FinishJob*)(d + st.off2), st.count, s);
                break;
            case S_FINISH: launch_finish((const FinishJob*)x, st.count, d_pc, log_n, s); break;
            case S_NTT_COLS: ntt_fast_forward_cols(ntt, (const LimbRef*)x, st.count, s); break;
            case S_NTT_ROWS: ntt_fast_forward_rows(ntt, (const LimbRef*)x, st.count, s); break;
            case S_ROWS_IP: rows_ip_launch(ntt, (const RowIpJob*)x, st.count, s); break;
            case S_ROWS_IP_FIN: rows_ip_fin_launch(ntt, (const RowIpFinJob*)x, st.count, s); break;
            case S_ADD_SUM:
                for (u32 i = 0; i < st.count; i += MAX_ROWS)
                    add_sum_launch((const MulSumJob*)x + i, (const u64* const*)(d + st.off2), std::min(MAX_ROWS, st.count - i),
                                   d_pc, log_n, s);
                break;
            case S_AUTOMORPH:
                for (u32 i = 0; i < st.count; i += MAX_ROWS)
                    automorphism_batch_launch((const AutJob*)x + i, std::min(MAX_ROWS, st.count - i), log_n, s);
                break;
            case S_MUL_SUM:
                for (u32 i = 0; i < st.count; i += MAX_ROWS)
                    mul_sum_launch((const MulSumJob*)x + i, (const MulSumTerm*)(d + st.off2), std::min(MAX_ROWS, st.count - i),
                                   d_pc, log_n, s);
                break;
        }
    }
}

void Context::execute(const Program& p, hipStream_t s) {
    if (p.steps.empty()) return;
    tick_launch();
    if (use_graphs && execute_graph(p, s)) return;
    BufPtr d = upload_blob(p.bytes.data(), p.bytes.size(), s);
    run(p, (const uint8_t*)d->ptr, s);
}

// Evicts plan i without waiting: its descriptor buffer returns to the pool in its stream's
// order (after the last replay), and the executable graph and event are destroyed once that
// replay has finished (retired until then, swept by later calls).
void Context::drop_plan(size_t i) {
    std::unique_ptr<Plan> pl = std::move(plans_[i]);
    plans_.erase(plans_.begin() + i);
    if (pl->desc) (void)hipFreeAsync(pl->desc, pl->s);
    pl->desc = nullptr;
    if (pl->done && hipEventQuery(pl->done) == hipErrorNotReady) {
        retired_.push_back(std::move(pl));
        return;
    }
    if (pl->done) (void)hipEventDestroy(pl->done);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
}

// Destroys the retired plans whose last replay has finished (wait: all of them, after it).
void Context::sweep_retired(bool wait) {
    for (size_t i = 0; i < retired_.size();) {
        Plan& pl = *retired_[i];
        if (wait) (void)hipEventSynchronize(pl.done);
        if (wait || hipEventQuery(pl.done) != hipErrorNotReady) {
            (void)hipEventDestroy(pl.done);
            if (pl.exec) (void)hipGraphExecDestroy(pl.exec);
            retired_.erase(retired_.begin() + i);
        } else {
            ++i;
        }
    }
}

void Context::clear_plans() {
    std::lock_guard<std::mutex> lk(plan_mu_);
    while (!plans_.empty()) drop_plan(plans_.size() - 1);
    sweep_retired(true);
}

// Returns false when the program ran on no graph (first sighting of its shape: the caller runs
// it directly). The graph holds only kernel launches; descriptors travel in the plan's buffer.
bool Context::execute_graph(const Program& p, hipStream_t s) {
    std::lock_guard<std::mutex> lk(plan_mu_);
    if (!retired_.empty()) sweep_retired(false);
    Plan* pl = nullptr;
    for (auto& q : plans_)
        if (q->s == s && q->bytes == p.bytes.size() && q->steps == p.steps) {
            pl = q.get();
            break;
        }
    if (!pl) {
        if (plans_.size() >= std::max<u32>(graph_cache_max, 1)) {
            size_t lru = 0;
            for (size_t i = 1; i < plans_.size(); ++i)
                if (plans_[i]->last_use < plans_[lru]->last_use) lru = i;
            drop_plan(lru);
        }
        auto np_ = std::make_unique<Plan>();
        np_->s = s;
        np_->steps = p.steps;
        np_->bytes = p.bytes.size();
        np_->last_use = ++plan_clock_;
        plans_.push_back(std::move(np_));
        return false;
    }
    pl->last_use = ++plan_clock_;
    if (!pl->desc) {
        hipError_t e = hipMallocAsync(&pl->desc, std::max<size_t>(p.bytes.size(), 16), s);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            pl->desc = nullptr;
            return false;
        }
    }
    // The previous replay (on this same stream: plans are per stream) reads the descriptor buffer
    // being rewritten. The rewrite is a DMA from pinned staging memory queued on s, so it runs
    // after that replay in stream order; the host does not wait.
    h2d(pl->desc, p.bytes.data(), p.bytes.size(), s);
    if (!pl->exec) {
        // Captured on the context's private stream (under plan_mu_, so no other thread touches
        // it): work that other threads queue on s meanwhile cannot leak into the graph, and a
        // synchronization of s elsewhere cannot end the capture. The graph is launched on s.
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        run(p, (const uint8_t*)pl->desc, capture_stream_);
        const hipError_t e_end = hipStreamEndCapture(capture_stream_, &g);
        hip_check(e_end, "hipStreamEndCapture");
        const hipError_t e_inst = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e_inst != hipSuccess) pl->exec = nullptr;
        hip_check(e_inst, "hipGraphInstantiate");
        ++graph_captures;
    }
    hip_check(hipGraphLaunch(pl->exec, s), "hipGraphLaunch");
    if (!pl->done) hip_check(hipEventCreateWithFlags(&pl->done, hipEventDisableTiming), "event");
    hip_check(hipEventRecord(pl->done, s), "event");
    ++graph_launches;
    return true;
}

void Context::ntt_poly(const u64* in, u64* out, u32 limbs, u32 level, u32 p_limbs, bool forward, hipStream_t s) {
    std::vector<LimbRef> refs;
    for (u32 i = 0; i < limbs; ++i) {
        u32 prime = i <= level ? i : nq + (i - level - 1);
        refs.push_back({in + i * n, out + i * n, prime, 0});
    }
    (void)p_limbs;
    Program p;
    p.step_ntt(forward ? S_NTT_FWD : S_NTT_INV, p.add(refs), (u32)refs.size(), 0, fused_usable());
    execute(p, s);
}

// keyswitch::decompose (keyswitch.rs:82-124) for a batch: INTT of every input limb, the
// centered exact conversion of every digit to its targets, the digit limbs copied, then the
// NTT of every converted limb (lift_digit_exact, keyswitch.rs:140-210).
bool Context::fuse_ip_possible() const {
    if (!ntt_fast_supported(log_n)) return false;
    for (u32 i = 0; i < nq; ++i)
        if (primes[i] >= FP_PRIME_BOUND) return false;
    return true;
}

// Limb descriptors grouped by prime (stable): a pass over many limbs then works on one prime at a
// time, and the integer path's row-pass twiddle table (1 MiB per prime at N = 2^16, read once per
// limb) stays in the L2 of the XCD running that prime instead of being fetched again per limb.
static void group_by_prime(std::vector<LimbRef>& r) {
    std::stable_sort(r.begin(), r.end(), [](const LimbRef& a, const LimbRef& b) { return a.prime < b.prime; });
}

bool Context::yform_usable(u32 mask_bit) const { return (conv_yform & mask_bit) && ntt_fast_supported(log_n); }

void Context::prog_decompose(Program& p, const std::vector<DecTask>& tasks, u32 variant) const {
    std::vector<LimbRef> intt, copies, nttr, cols;
    std::vector<ConvJob> jobs;
    u32 max_tgt = 0;
    const bool yf = yform_usable(variant == 2 && conv.mf_ok ? 4 : 1);
    for (const DecTask& t : tasks) {
        const u32 L = t.level + 1;
        const u64 ds = digit_words(t.level);
        auto dr = digit_ranges(t.level);
        for (u32 j = 0; j < dr.size(); ++j)
            for (u32 i = dr[j].first; i < dr[j].second; ++i)
                intt.push_back({t.d + i * n, t.x + i * n, i,
                                yf ? conv_src_first_[modup_plan_[t.level][j]] + (i - dr[j].first) : 0});
        for (u32 j = 0; j < dr.size(); ++j) {
            u64* dj = t.dig + j * ds;
            jobs.push_back({t.x, dj, modup_plan_[t.level][j], 0});
            if (t.copy_own)
                for (u32 i = dr[j].first; i < dr[j].second; ++i) copies.push_back({t.d + i * n, dj + i * n, i, 0});
            for (u32 i = 0; i < L; ++i)
                if (i < dr[j].first || i >= dr[j].second) (t.fuse ? cols : nttr).push_back({dj + i * n, dj + i * n, i, 0});
            // P targets are transformed whole (their inner product is not fused with a row pass)
            for (u32 k = 0; k < np; ++k) nttr.push_back({dj + (L + k) * n, dj + (L + k) * n, nq + k, 0});
        }
        max_tgt = std::max(max_tgt, modup_maxtgt_[t.level]);
    }
    group_by_prime(nttr);
    if (yf)
        p.step(S_NTT_INV_Y, p.add(intt), (u32)intt.size());
    else
        p.step_ntt(S_NTT_INV, p.add(intt), (u32)intt.size(), 0, fused_usable());
    p.step(S_CONV, p.add(jobs), (u32)jobs.size(), 0, max_tgt, alpha);
    if (!jobs.empty()) p.steps.back().yform = yf;
    p.step(S_COPY, p.add(copies), (u32)copies.size());
    p.step_ntt(S_NTT_FWD, p.add(nttr), (u32)nttr.size(), 0, fused_usable());
    p.step(S_NTT_COLS, p.add(cols), (u32)cols.size());
}

// keyswitch::apply_key (keyswitch.rs:230-309) for a batch: the inner product over Q_l * P, then
// mod_down of both sums.
void Context::prog_apply(Program& p, const std::vector<ApplyTask>& tasks) const {
    std::vector<IpJob> ip;
    u32 max_l = 0;
    for (const ApplyTask& t : tasks) max_l = std::max(max_l, t.level + 1);
    // Inner-product rows ordered by key limb across tasks, so rows reading the same key limb
    // run back to back (the shared key is re-read from L2, not HBM).
    for (u32 k = 0; k < np; ++k)
        for (const ApplyTask& t : tasks) {
            const u32 L = t.level + 1, li = L + k;
            const u64 aw = (u64)(L + np) * n;
            ip.push_back({t.dig + li * n, digit_words(t.level), t.key->buf->ptr + (t.key->max_level + 1 + k) * n,
                          t.key->poly_words(), t.acc + li * n, t.acc + aw + li * n, t.galois, nullptr,
                          ndigits(t.level), nq + k, 0, 0});
        }
    bool all_fused = !tasks.empty();
    for (const ApplyTask& t : tasks) all_fused = all_fused && t.fused;
    if (all_fused && fuse_fin) {
        // The P limbs' inner product, mod_down of the P parts up to the column pass of the
        // converted limbs, then per Q limb one kernel: row pass + inner product + finish.
        p.step(S_IP, p.add(ip), (u32)ip.size());
        std::vector<LimbRef> intt, cols;
        std::vector<ConvJob> jobs;
        const bool yf = yform_usable(2);
        for (const ApplyTask& t : tasks) {
            const u32 L = t.level + 1;
            const u64 aw = (u64)(L + np) * n;
            for (u32 part = 0; part < 2; ++part) {
                const u64* acc = t.acc + part * aw;
                u64* y = t.y + (u64)part * np * n;
                u64* z = t.z + (u64)part * L * n;
                const u32 sf = conv_src_first_[moddown_plan_[t.level]];
                for (u32 k = 0; k < np; ++k) intt.push_back({acc + (L + k) * n, y + k * n, nq + k, yf ? sf + k : 0});
                jobs.push_back({y, z, moddown_plan_[t.level], 0});
                for (u32 i = 0; i < L; ++i) cols.push_back({z + i * n, z + i * n, i, 0});
            }
        }
        std::vector<RowIpFinJob> rf;
        for (u32 i = 0; i < max_l; ++i)
            for (const ApplyTask& t : tasks) {
                const u32 L = t.level + 1;
                if (i >= L) continue;
                RowIpFinJob j{};
                j.ip = {t.dig + i * n, digit_words(t.level), t.own + i * n, t.key->buf->ptr + i * n, t.key->poly_words(),
                        nullptr, nullptr, ndigits(t.level), i, i / alpha, 0};
                j.zb = t.z + i * n;
                j.za = t.z + ((u64)L + i) * n;
                j.out_b = t.out_b + i * n;
                j.out_a = t.out_a + i * n;
                j.add_b = t.add_b ? t.add_b + i * n : nullptr;
                j.add_a = t.add_a ? t.add_a + i * n : nullptr;
                j.inv = {(double)pinv_[i], (double)pinv_[i] / (double)primes[i]};
                rf.push_back(j);
            }
        group_by_prime(intt);
        if (yf)
            p.step(S_NTT_INV_Y, p.add(intt), (u32)intt.size());
        else
            p.step_ntt(S_NTT_INV, p.add(intt), (u32)intt.size(), 0, fused_usable());
        p.step(S_CONV, p.add(jobs), (u32)jobs.size(), 0, max_l, np);
        if (!jobs.empty()) p.steps.back().yform = yf;
        if (conv_variant_md < 3 && !jobs.empty()) p.steps.back().sched = 1u | (conv_variant_md << 1) | (conv_tg << 8);
        p.step(S_NTT_COLS, p.add(cols), (u32)cols.size());
        p.step(S_ROWS_IP_FIN, p.add(rf), (u32)rf.size());
        return;
    }
    std::vector<RowIpJob> rip;
    for (u32 i = 0; i < max_l; ++i)
        for (const ApplyTask& t : tasks) {
            const u32 L = t.level + 1;
            if (i >= L) continue;
            const u64 aw = (u64)(L + np) * n;
            if (t.fused) {
                rip.push_back({t.dig + i * n, digit_words(t.level), t.own + i * n, t.key->buf->ptr + i * n,
                               t.key->poly_words(), t.acc + i * n, t.acc + aw + i * n, ndigits(t.level), i,
                               i / alpha, 0});
                continue;
            }
            ip.push_back({t.dig + i * n, digit_words(t.level), t.key->buf->ptr + i * n, t.key->poly_words(),
                          t.acc + i * n, t.acc + aw + i * n, t.galois, t.own ? t.own + i * n : nullptr,
                          ndigits(t.level), i, i / alpha, 0});
        }
    p.step(S_IP, p.add(ip), (u32)ip.size());
    p.step(S_ROWS_IP, p.add(rip), (u32)rip.size());
    std::vector<MdTask> md;
    for (const ApplyTask& t : tasks) {
        const u32 L = t.level + 1;
        const u64 aw = (u64)(L + np) * n;
        md.push_back({t.acc, t.level, t.out_b, t.add_b, t.y, t.z});
        md.push_back({t.acc + aw, t.level, t.out_a, t.add_a, t.y + (u64)np * n, t.z + (u64)L * n});
    }
    prog_mod_down(p, md);
}

// RnsPoly::mod_down (baseconv.rs:459-511) for a batch: INTT of the P limbs, the centered exact
// conversion P -> Q_l, NTT, and (x - t) * P^-1 (+ the optional addend).
void Context::prog_mod_down(Program& p, const std::vector<MdTask>& tasks) const {
    std::vector<LimbRef> intt, nttr;
    std::vector<ConvJob> jobs;
    std::vector<FinishJob> fin;
    u32 max_l = 0;
    const bool yf = yform_usable(2);
    for (const MdTask& t : tasks) {
        const u32 L = t.level + 1;
        max_l = std::max(max_l, L);
        const u32 sf = conv_src_first_[moddown_plan_[t.level]];
        for (u32 k = 0; k < np; ++k) intt.push_back({t.acc + (L + k) * n, t.y + k * n, nq + k, yf ? sf + k : 0});
        jobs.push_back({t.y, t.z, moddown_plan_[t.level], 0});
        for (u32 i = 0; i < L; ++i) {
            u64* zi = t.z + i * n;
            nttr.push_back({zi, zi, i, 0});
            fin.push_back({t.acc + i * n, zi, t.out + i * n, t.add ? t.add + i * n : nullptr, pinv_[i], pinv_s_[i], i, 0,
                           {(double)pinv_[i], (double)pinv_[i] / (double)primes[i]}});
        }
    }
    group_by_prime(intt);
    if (yf)
        p.step(S_NTT_INV_Y, p.add(intt), (u32)intt.size());
    else
        p.step_ntt(S_NTT_INV, p.add(intt), (u32)intt.size(), 0, fused_usable());
    p.step(S_CONV, p.add(jobs), (u32)jobs.size(), 0, max_l, np);
    if (!jobs.empty()) p.steps.back().yform = yf;
    if (conv_variant_md < 3 && !jobs.empty()) p.steps.back().sched = 1u | (conv_variant_md << 1) | (conv_tg << 8);
    const size_t o_ntt = p.add(nttr), o_fin = p.add(fin);
    p.step_ntt(S_NTT_FWD_FIN, o_ntt, (u32)nttr.size(), o_fin, fused_usable());
}

// RnsPoly::rescale_by_last in NTT form (baseconv.rs:258-291) for a batch.
void Context::prog_rescale(Program& p, const std::vector<RescaleTask>& tasks) const {
    std::vector<LimbRef> intt, nttr;
    std::vector<ConvJob> jobs;
    std::vector<FinishJob> fin;
    u32 max_t = 0;
    for (const RescaleTask& t : tasks) {
        const u32 l = t.level;
        intt.push_back({t.in + l * n, t.x, l, 0});
        jobs.push_back({t.x, t.z, rescale_plan_[l], 0});
        for (u32 i = 0; i < l; ++i) {
            nttr.push_back({t.z + i * n, t.z + i * n, i, 0});
            fin.push_back({t.in + i * n, t.z + i * n, t.out + i * n, nullptr, qlinv_[l][i], qlinv_s_[l][i], i, 0,
                           {(double)qlinv_[l][i], (double)qlinv_[l][i] / (double)primes[i]}});
        }
        max_t = std::max(max_t, l);
    }
    p.step_ntt(S_NTT_INV, p.add(intt), (u32)intt.size(), 0, fused_usable());
    p.step(S_CONV, p.add(jobs), (u32)jobs.size(), 0, max_t, 1);
    const size_t o_ntt = p.add(nttr), o_fin = p.add(fin);
    p.step_ntt(S_NTT_FWD_FIN, o_ntt, (u32)nttr.size(), o_fin, fused_usable());
}

// RnsPoly::raise_from_level0 (crates/ckks-train-core/src/poly.rs:298-345; the device form of
// crates/ckks-train-cuda/src/basis.rs:513-533): limb 0 copied as it is; its coefficients (after an
// inverse NTT for NTT-form input) converted to q_1..q_top by the single-source centered lift,
// then transformed forward again.
void Context::prog_raise(Program& p, const std::vector<RaiseTask>& tasks) const {
    std::vector<LimbRef> copy, intt, nttr;
    std::vector<ConvJob> jobs;
    u32 max_t = 0;
    for (const RaiseTask& t : tasks) {
        if (t.top >= nq) fail(CKH_ERR_INVALID_PARAMS, "level out of range");
        copy.push_back({t.in, t.out, 0, 0});
        if (t.top == 0) continue;
        const u64* src = t.in;
        if (t.ntt) {
            intt.push_back({t.in, t.x, 0, 0});
            src = t.x;
            for (u32 i = 1; i <= t.top; ++i) nttr.push_back({t.out + i * n, t.out + i * n, i, 0});
        }
        jobs.push_back({src, t.out, raise_plan_[t.top], 0});
        max_t = std::max(max_t, t.top);
    }
    p.step(S_COPY, p.add(copy), (u32)copy.size());
    p.step_ntt(S_NTT_INV, p.add(intt), (u32)intt.size(), 0, fused_usable());
    p.step(S_CONV, p.add(jobs), (u32)jobs.size(), 0, max_t, 1);
    p.step_ntt(S_NTT_FWD, p.add(nttr), (u32)nttr.size(), 0, fused_usable());
}

void Context::raise_level0(const std::vector<RaiseTask>& tasks, hipStream_t s) {
    Program p;
    prog_raise(p, tasks);
    execute(p, s);
}

void Context::decompose(const std::vector<DecTask>& tasks, hipStream_t s) {
    Program p;
    prog_decompose(p, tasks, conv_variant);
    execute(p, s);
}

void Context::apply(const std::vector<ApplyTask>& tasks, hipStream_t s) {
    Program p;
    prog_apply(p, tasks);
    execute(p, s);
}

void Context::mod_down(const std::vector<MdTask>& tasks, hipStream_t s) {
    Program p;
    prog_mod_down(p, tasks);
    execute(p, s);
}

void Context::rescale(const std::vector<RescaleTask>& tasks, hipStream_t s) {
    Program p;
    prog_rescale(p, tasks);
    execute(p, s);
}

// NTT of the converted limbs followed by the finish; fused into the row pass where the
// register-radix NTT applies.
void Context::launch_ntt_finish(const LimbRef* refs, const FinishJob* fin, u32 n_refs, hipStream_t s) const {
    if (ntt_fast_supported(log_n)) {
        const u32 chunk = (ntt_chunk_fwd == 0 || ntt_chunk_fwd > MAX_ROWS) ? MAX_ROWS : ntt_chunk_fwd;
        for (u32 i = 0; i < n_refs; i += chunk)
            ntt_fast_forward_finish(ntt, refs + i, fin + i, std::min(chunk, n_refs - i), s);
        return;
    }
    launch_ntt(*this, refs, n_refs, true, s);
    launch_finish(fin, n_refs, d_pc, log_n, s);
}

}  // namespace ckh
