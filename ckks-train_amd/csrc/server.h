// Serving loop: keys resident, requests admitted by HBM footprint, continuous batching on
// several streams, per-request latency.
#pragma once

#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "evaluator.h"

namespace ckh {

struct ServedOp {
    u64 ticket = 0;
    u32 kind = 0;
    std::shared_ptr<Evaluator> ev;  // kept alive (with its keys) until the request completes
    CtPtr a, b;
    std::vector<int32_t> steps;
    bool inject_failure = false;
    double t_arrival = 0, t_done = 0;
    u64 footprint = 0;
    hipEvent_t ready = nullptr;  // inputs ready on the submitter's stream
    ckh_status status = CKH_OK;
    std::string err;
    std::vector<CtPtr> outs;
    bool done = false;
    bool launched = false;  // its batch is in flight (completion marks it done)
};
using OpPtr = std::shared_ptr<ServedOp>;

class Server {
   public:
    explicit Server(const ckh_server_config& cfg);
    ~Server();
    double now() const;
    // op carries kind, evaluator, inputs, steps, arrival and the failure hook.
    u64 submit(const OpPtr& op);
    // true when the request has finished; fills op (outputs, status, latency).
    bool poll(u64 ticket, bool wait, OpPtr& op);
    ckh_server_stats stats();

   private:
    struct Batch {
        Context* ctx;
        hipEvent_t done;
        std::vector<OpPtr> ops;
    };
    void loop();
    void launch(std::vector<OpPtr>& ops);
    void fail_op(const OpPtr& op, ckh_status st, const std::string& msg);
    void fail_admitted(std::vector<OpPtr>& admitted, ckh_status st, const std::string& msg);
    u64 footprint(const ServedOp& op) const;

    ckh_server_config cfg_;
    std::chrono::steady_clock::time_point t0_;
    std::thread th_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<OpPtr> pending_;
    std::vector<Batch> inflight_;
    std::unordered_map<u64, OpPtr> table_;
    std::unordered_map<const Context*, u32> rr_;
    u64 next_ticket_ = 1;
    u64 in_use_ = 0, peak_ = 0;
    u64 submitted_ = 0, completed_ = 0, failed_ = 0, rejected_ = 0, batches_ = 0;
    double launch_ms_ = 0, launch_max_ms_ = 0;
    double first_arrival_ = -1, last_done_ = 0;
    std::vector<double> lat_ms_;
    bool stop_ = false;
};

}  // namespace ckh
