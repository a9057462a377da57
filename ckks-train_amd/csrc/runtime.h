// Host runtime of the gfx950 RNS-CKKS library: device context, tables, stream-ordered memory,
// and the key-switching engine (decompose -> inner product -> mod_down) over batches of tasks.
#pragma once

#include <atomic>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ckh.h"
#include "basis.h"
#include "common.h"
#include "lifetime.h"
#include "ntt.h"
#include "ops.h"

namespace ckh {

// Internal error, mapped to a ckh_status at the C boundary.
struct Error {
    ckh_status status;
    std::string msg;
};

[[noreturn]] void fail(ckh_status st, const std::string& msg);
void hip_check(hipError_t e, const char* what);

class Context;


// A device allocation from the context's block cache (or the default pool); released in its
// stream's order when the last reference goes.
struct Buffer {
    Context* ctx;
    std::shared_ptr<Context> keep;  // a buffer keeps its context (and streams) alive
    u64* ptr;
    size_t words;
    hipStream_t stream;
    size_t cls = 0;  // size class of a block from the context's cache (0: from the default pool)
    // The other streams that have had work on this buffer queued (a caller's input read on the
    // serving loop's stream, a key read by every thread's stream): the free waits for each.
    std::mutex use_mu;
    std::vector<hipStream_t> users;
    void note_use(hipStream_t s) {
        if (s == stream) return;
        std::lock_guard<std::mutex> lk(use_mu);
        for (hipStream_t u : users)
            if (u == s) return;
        users.push_back(s);
    }
    ~Buffer();
};
using BufPtr = std::shared_ptr<Buffer>;

// A polynomial: Q limbs 0..=level, then p_limbs P limbs, inside a buffer.
struct Poly {
    BufPtr buf;
    u64 off;  // word offset in buf
    u32 level;
    u32 p_limbs;
    bool ntt;
    u64* data() const { return buf->ptr + off; }
    u32 limbs() const { return level + 1 + p_limbs; }
};
using PolyPtr = std::shared_ptr<Poly>;

// A key-switching key in the reference's device order (parts x {b, a} x (kml + 1 + k) limbs).
struct Ksk {
    BufPtr buf;
    u32 max_level;
    u32 parts;
    u64 galois;  // 0 = relinearization
    Context* ctx;
    u64 poly_words() const;
};
using KskPtr = std::shared_ptr<Ksk>;

struct Ct {
    std::vector<PolyPtr> comps;
    double scale;
    u32 level() const { return comps.empty() ? 0 : comps[0]->level; }
};
using CtPtr = std::shared_ptr<Ct>;

// One decomposition: NTT-form d (level l) -> digits (ndig QpPolys of l + 1 + k limbs, stride
// ds words) using x (l + 1 limbs) as scratch.
struct DecTask {
    const u64* d;
    u32 level;
    u64* dig;
    u64* x;
    bool copy_own = true;  // false: the digits' own limbs stay in d (an ApplyTask reads them there)
    bool fuse = false;     // Q-prime targets get only the column pass (an ApplyTask with fused set
                           // finishes them inside its inner product)
};

// One inner product + mod_down: digits (from a DecTask) times key (galois != 0: digits read
// through sigma_galois) -> (out_b, out_a) at level l, plus optional addends. acc (2 (l+1+k)
// limbs), y (2k limbs) and z (2 (l+1) limbs) are scratch.
struct ApplyTask {
    const u64* dig;
    u32 level;
    u64 galois;
    const Ksk* key;
    u64* out_b;
    u64* out_a;
    const u64* add_b;
    const u64* add_a;
    u64* acc;
    u64* y;
    u64* z;
    const u64* own = nullptr;  // set: the digits' own limbs are read from this d (not copied)
    bool fused = false;        // the digits come from a DecTask with fuse set (needs own)
};

// One mod_down of a QpPoly sum (level l, l + 1 + k limbs, NTT form) -> out (l + 1 limbs) plus an
// optional addend; y (k limbs) and z (l + 1 limbs) are scratch.
struct MdTask {
    const u64* acc;
    u32 level;
    u64* out;
    const u64* add;
    u64* y;
    u64* z;
};

// One rescale of a polynomial (level l -> l - 1): out (l limbs), scratch x (1 limb), z (l limbs).
struct RescaleTask {
    const u64* in;
    u32 level;
    u64* out;
    u64* x;
    u64* z;
};

// One raise from level 0 (RnsPoly::raise_from_level0): limb 0 of in (NTT form when ntt) copied
// to out, limbs 1..=top of out the centered lift of its coefficients; scratch x (1 limb).
struct RaiseTask {
    const u64* in;
    u32 top;
    u64* out;
    u64* x;
    bool ntt;
};

// Descriptor bytes plus the launch steps that read them (see runtime.cpp, "engine").
enum StepKind : uint8_t {
    S_NTT_FWD, S_NTT_INV, S_CONV, S_COPY, S_IP, S_NTT_FWD_FIN, S_FINISH, S_NTT_COLS, S_NTT_ROWS, S_ROWS_IP,
    S_MUL_SUM, S_AUTOMORPH, S_ADD_SUM, S_ROWS_IP_FIN, S_NTT_INV_Y
};
struct Program {
    struct Step {
        StepKind kind;
        u32 count;
        size_t off, off2;
        u32 a, b;
        u32 sched = 0;  // S_CONV: 0 = the context's schedule, else 1 | variant << 1 | tg << 8
        size_t ctl = 0;  // NTT steps: offset of the one-launch transform's zeroed control block
        bool yform = false;  // S_CONV: the inputs are in y form (an S_NTT_INV_Y step wrote them)
        bool operator==(const Step& o) const {
            return kind == o.kind && count == o.count && off == o.off && off2 == o.off2 && a == o.a && b == o.b &&
                   sched == o.sched && ctl == o.ctl && yform == o.yform;
        }
    };
    std::vector<uint8_t> bytes;
    std::vector<Step> steps;
    // Appends v at a 16-byte aligned offset of the descriptor bytes and returns the offset.
    template <class T>
    size_t add(const std::vector<T>& v) {
        const size_t off = (bytes.size() + 15) & ~(size_t)15;
        bytes.resize(off + v.size() * sizeof(T));
        if (!v.empty()) std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
    void step(StepKind k, size_t off, u32 count, size_t off2 = 0, u32 a = 0, u32 b = 0);
    // An NTT step (S_NTT_FWD, S_NTT_INV, or S_NTT_FWD_FIN with its finish jobs at off2); with
    // one_launch, a zeroed control block in the descriptor bytes for the one-launch transforms
    // (every execution uploads the bytes again, so each launch finds its block zero).
    void step_ntt(StepKind k, size_t off, u32 count, size_t off2, bool one_launch);
    // Every base conversion of the program with this kernel variant and targets per workgroup.
    void set_conv_schedule(u32 variant, u32 tg) {
        for (Step& st : steps)
            if (st.kind == S_CONV) st.sched = 1u | (variant << 1) | (tg << 8);
    }
};

struct HostConvTables;

class Context : public std::enable_shared_from_this<Context> {
   public:
    Context(int ordinal, const ckh_params& p, u32 n_streams);
    ~Context();

    int device = 0;
    u32 log_n = 0;
    u64 n = 0;
    u32 nq = 0, np = 0, dnum = 0, alpha = 0;
    std::vector<u64> primes;  // Q then P

    NttTables ntt{};
    ConvTables conv{};
    PrimeConst* d_pc = nullptr;

    // Stream of the calling thread: one of the n_streams() caller streams (ckh_set_stream) or,
    // for the serving loop's thread, one of SERVICE_STREAMS streams no caller can select (a
    // graph is captured on a stream only by the thread working on it).
    static constexpr u32 SERVICE_STREAMS = 4;
    hipStream_t stream() const;
    void set_stream(u32 index);
    void set_service_stream(u32 k);  // the k-th of service_streams (tuning key, 1..SERVICE_STREAMS)
    u32 service_streams = SERVICE_STREAMS;
    u32 n_streams() const { return n_user_streams_; }
    u32 n_all_streams() const { return (u32)streams_.size(); }
    // Waits for every stream, the serving loop's included, outside any graph capture (captures
    // hold the plan lock; synchronizing a capturing stream would end its capture).
    void synchronize_all();
    hipStream_t stream_at(u32 i) const { return streams_[i]; }

    // Device memory is stream-ordered: a buffer is freed on the stream it was allocated on, after
    // the allocation stream has been made to wait for the work enqueued so far on the releasing
    // thread's stream and on every other stream that used the buffer (Buffer::users).
    // Tuning key "pool": 1 (default) = the context's block cache: device blocks in size classes
    // (four per power of two), kept for reuse instead of returned to the driver (whose pools
    // released memory at synchronizations and stalled the host for the whole device queue when
    // growing under load); a freed block serves its own stream at once (stream order) and any
    // other stream once an event recorded after its last use has completed; 0 = the device's
    // default stream-ordered pool with the runtime's reuse rules (which may hand a block to
    // another stream before its free has run, behind a dependency the runtime inserts).
    BufPtr alloc(size_t words, hipStream_t s);
    BufPtr alloc(size_t words) { return alloc(words, stream()); }
    void free_buffer(Buffer* b);
    u32 pool_mode = 1;
    // The pool's bytes in use, its release threshold (bytes kept reserved after a
    // synchronization; default: all) and a trim to keep_bytes after every stream has finished
    // (crates/ckks-train-cuda/src/device.rs:478-527).
    u64 pool_used_memory();
    u64 pool_unused_reserve();  // reserved by the pool but not in use
    void set_pool_release_threshold(u64 bytes);
    void trim_pool(u64 keep_bytes);
    // Test hook (tuning key "fail_alloc_after"): the n-th allocation from now fails with a device
    // error, as an out-of-memory would; negative = off.
    std::atomic<int64_t> fail_alloc_after{-1};
    // Test hooks as the reference's failing backend (crates/ckks-train-cuda/tests/cases/
    // failing.rs:34-67): "fail_launch_after" makes the n-th primitive launch from now (a
    // program's launches, or one direct kernel launch) fail with a device error before anything
    // of it is queued, and disarms itself; "fail_upload" (1) makes every upload of host words
    // fail so.
    std::atomic<int64_t> fail_launch_after{-1};
    std::atomic<bool> fail_upload{false};
    void tick_launch() {
        if (fail_launch_after.load() >= 0 && fail_launch_after.fetch_sub(1) == 0)
            fail(CKH_ERR_DEVICE, "injected launch failure");
    }
    void tick_upload() const {
        if (fail_upload.load()) fail(CKH_ERR_DEVICE, "injected upload failure");
    }
    // Block-cache statistics (read-only tuning keys "stat_mallocs", "stat_malloc_us", "stat_oom"):
    // blocks obtained from the driver, the host time spent doing so, and allocation failures.
    std::atomic<u64> stat_mallocs{0}, stat_malloc_us{0}, stat_oom{0};
    std::atomic<u64> live_bytes{0};

    // Digit ranges of a level-l polynomial (keyswitch.rs:68-74).
    std::vector<std::pair<u32, u32>> digit_ranges(u32 level) const;
    u32 ndigits(u32 level) const;
    u64 digit_words(u32 level) const { return (u64)(level + 1 + np) * n; }

    // The engine: build programs, run them (descriptors at d), or both (execute).
    // variant: the conversion kernel variant the program's ModUp conversion will run with (it
    // decides whether the inverse NTT leaves the y form, tuning key "conv_yform").
    void prog_decompose(Program& p, const std::vector<DecTask>& tasks, u32 variant) const;
    void prog_apply(Program& p, const std::vector<ApplyTask>& tasks) const;
    void prog_mod_down(Program& p, const std::vector<MdTask>& tasks) const;
    void prog_rescale(Program& p, const std::vector<RescaleTask>& tasks) const;
    void run(const Program& p, const uint8_t* d, hipStream_t s) const;
    void execute(const Program& p, hipStream_t s);
    void decompose(const std::vector<DecTask>& tasks, hipStream_t s);
    void apply(const std::vector<ApplyTask>& tasks, hipStream_t s);
    void mod_down(const std::vector<MdTask>& tasks, hipStream_t s);
    void rescale(const std::vector<RescaleTask>& tasks, hipStream_t s);
    void prog_raise(Program& p, const std::vector<RaiseTask>& tasks) const;
    void raise_level0(const std::vector<RaiseTask>& tasks, hipStream_t s);
    void ntt_poly(const u64* in, u64* out, u32 limbs, u32 level, u32 p_limbs, bool forward, hipStream_t s);

    void launch_ntt_finish(const LimbRef* refs, const FinishJob* fin, u32 n_refs, hipStream_t s) const;

    // Tuning (autotune.cpp): limbs per column-pass/row-pass launch pair (0 = all at once). A
    // chunk small enough keeps the pass intermediate in the on-die caches between the passes.
    u32 ntt_chunk_fwd = 0, ntt_chunk_inv = 0;
    u32 conv_variant = 0, conv_tg = 8;  // base-conversion kernel variant and targets per workgroup
    // The variant of mod_down's conversions (P -> Q) in batched programs (tuning key
    // "conv_variant_md"; 3 = conv_variant). Default 1: with 60-bit sources the exact f64 piece
    // sums beat two modular products per source (step 19.98 -> 19.54 ms on the 256-switch
    // workload), while the ModUp conversions (50-bit sources) keep conv_variant.
    u32 conv_variant_md = 1;
    // Batched key switches of at most small_limbs input limbs in all (a single switch, a few
    // small ones) have too few workgroups per launch to fill the device; they run their own
    // schedule (tuning keys "small_limbs", "conv_variant_small", "conv_tg_small", "fuse_ip_small").
    u32 small_limbs = 48, conv_variant_s = 2, conv_tg_s = 8, fuse_ip_s = 0;
    std::string autotune();
    // The key-switch candidates of autotune() timed on a caller's batch (autotune.cpp).
    std::string autotune_ks(const std::vector<const Poly*>& ds, const Ksk& key);
    std::string autotune_small(const Poly& d, const Ksk& key);
    void set_tuning(const std::string& key, int64_t value);
    int64_t get_tuning(const std::string& key) const;
    // This GPU's fixed workspace (ckh_key_switch_batch_static), looked up at context creation.
    uint8_t* static_scratch = nullptr;
    // Copies a host blob to a fresh device buffer on stream s.
    BufPtr upload_blob(const void* host, size_t bytes, hipStream_t s);
    // Host-to-device copy ordered on stream s: the bytes are staged in a pinned ring (so the copy
    // is an ordinary stream-ordered DMA and `src` may be reused on return). Ring slots are reused
    // only after the copies that read them have completed.
    void h2d(void* dst, const void* src, size_t bytes, hipStream_t s);

    // Captured programs (hipGraph): a program whose launch steps (kinds, grid sizes, descriptor
    // offsets) were seen before on the same stream is captured once and then replayed, with its
    // descriptor bytes copied into the plan's own device buffer before each replay. Tuning key
    // "graphs" (1 = on) and "graph_cache" (plans kept per context, least recently used evicted).
    u32 use_graphs = 1, graph_cache_max = 64;
    // The one-launch NTTs (tuning key "ntt_fused", 1 = on where the device's XCD layout was
    // verified at creation, off by default: measured no faster than the two launches on the key
    // switch, slower on small launches; "fused_lag": slots between a limb's first- and
    // second-pass tiles; "fused_wg": workgroups per CU).
    FusedCfg fused{};
    u32 ntt_fused = 0;
    u32 cu_count = 0;
    // Hoisted-rotation requests the serving loop switches in one program (tuning key
    // "hoisted_batch"; 1 = one request per program).
    u32 hoisted_batch = 4;
    bool fused_usable() const { return ntt_fused && fused.nx && ntt_fast_supported(log_n); }
    u32* d_err = nullptr;  // device error word (a one-launch transform's wait overran its bound)
    // Raises a Device error (and clears the word) if a kernel has set the device error word.
    void check_device_error(hipStream_t s);
    // Row pass fused into the inner product of batched key switches (tuning key "fuse_ip");
    // possible when the register-radix NTT applies and every Q prime is below 2^50. With it,
    // mod_down's finish of the Q limbs in the same kernel (tuning key "fuse_fin"): the sums never
    // reach memory.
    u32 fuse_ip = 1, fuse_fin = 1;
    bool fuse_ip_possible() const;
    // The inverse NTT before a ModUp or mod_down conversion leaves its limbs in the conversion's
    // y form (YForm; step S_NTT_INV_Y), when the register-radix NTT applies: the conversion then
    // skips its per-source product. Tuning key "conv_yform", a mask: 1 = ModUp conversions on the
    // vector kernel, 2 = mod_down conversions, 4 = ModUp conversions on the matrix-core kernel.
    // Default 7 (every kind).
    u32 conv_yform = 7;
    bool yform_usable(u32 mask_bit) const;
    std::atomic<u64> graph_captures{0}, graph_launches{0};

   private:
    struct Staged {
        size_t off, len;
        hipEvent_t ev;
    };
    std::mutex ring_mu_;
    uint8_t* ring_ = nullptr;  // pinned host memory
    size_t ring_cap_ = 0, ring_head_ = 0;
    std::deque<Staged> ring_busy_;
    std::vector<hipEvent_t> ring_events_;
    size_t ring_reserve(size_t len);
    struct Plan {
        hipStream_t s = nullptr;
        std::vector<Program::Step> steps;
        size_t bytes = 0;
        void* desc = nullptr;  // stream-ordered allocation on s, outside Buffer (which would keep the
                               // context alive)
        hipGraphExec_t exec = nullptr;
        hipEvent_t done = nullptr;  // recorded after each replay
        u64 last_use = 0;
    };
    std::mutex plan_mu_;
    hipStream_t capture_stream_ = nullptr;  // graphs are captured here (under plan_mu_), launched on the caller's
    std::vector<std::unique_ptr<Plan>> plans_;
    std::vector<std::unique_ptr<Plan>> retired_;  // evicted plans whose last replay may still run
    void sweep_retired(bool wait);
    u64 plan_clock_ = 0;
    bool execute_graph(const Program& p, hipStream_t s);
    void drop_plan(size_t i);

   public:
    // Drops every captured plan (a tuning change alters the launch sequences).
    void clear_plans();

    std::vector<hipStream_t> streams_;  // caller streams, then the service streams
    // The block cache of pool mode 1: free blocks per size class, each with the stream of its
    // last use and an event recorded there after it.
    struct Cached {
        void* p;
        hipStream_t s;
        hipEvent_t ready;
    };
    std::mutex cache_mu_;
    std::map<size_t, std::vector<Cached>> cache_;
    std::vector<hipEvent_t> cache_events_;
    u64 cache_bytes_ = 0, cache_limit_ = ~0ull;
    // A free block of class cls or of a larger class up to 2 cls (cls is set to the block's).
    void* cache_take(size_t& cls, hipStream_t s);
    void cache_put(void* p, size_t cls, hipStream_t s);
    void cache_release(u64 keep_bytes);  // after every stream has finished
    u32 n_user_streams_ = 0;
    std::vector<void*> perm_allocs_;
    std::vector<std::vector<u32>> modup_plan_;  // [level][digit]
    std::vector<u32> modup_maxtgt_;            // [level]
    std::vector<u32> moddown_plan_;            // [level]
    std::vector<u32> conv_src_first_;          // [plan]: its first source in the ConvSrc table
    std::vector<u32> rescale_plan_;            // [level], level >= 1
    std::vector<u32> raise_plan_;              // [top], top >= 1: q_0 -> q_1..q_top
    std::vector<u64> pinv_, pinv_s_;           // P^-1 mod q_i
    std::vector<std::vector<u64>> qlinv_, qlinv_s_;  // [l][i]: q_l^-1 mod q_i
    void build_tables();
    void build_mf_tables(HostConvTables& ct);
    template <class T>
    T* dev_copy(const std::vector<T>& v);
};

// Launch helpers that split grids above the 65535 row limit.
void launch_ntt(const Context& c, const LimbRef* d_refs, u32 nrefs, bool forward, hipStream_t s);

// Device address of this GPU's fixed workspace (CKH_STATIC_SCRATCH_BYTES, static_scratch.hip),
// looked up once per device.
uint8_t* static_scratch_address(int device);

}  // namespace ckh
