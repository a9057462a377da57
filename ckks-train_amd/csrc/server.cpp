// Serving loop (see include/ckh.h, "serving").
#include "server.h"

#include <algorithm>
#include <cmath>

namespace ckh {

Server::Server(const ckh_server_config& cfg) : cfg_(cfg), t0_(std::chrono::steady_clock::now()) {
    if (cfg_.max_batch == 0) cfg_.max_batch = 64;
    if (cfg_.max_inflight == 0) cfg_.max_inflight = 4;
    th_ = std::thread([this] { loop(); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    th_.join();
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

// Device bytes of a request while it runs: outputs plus the scratch of its key switches.
u64 Server::footprint(const ServedOp& op) const {
    const Context& c = op.ev->ctx;
    const u64 L = op.a->level() + 1, k = c.np, d = c.ndigits(op.a->level());
    const u64 ks = L + d * (L + k) + 2 * (L + k) + 2 * k + 2 * L;  // one key switch, limbs
    u64 limbs = 0;
    switch (op.kind) {
        case CKH_OP_RELINEARIZE: limbs = ks + 2 * L; break;
        case CKH_OP_HOISTED_ROTATE:
            limbs = d * (L + k) + L + op.steps.size() * (2 * (L + k) + 2 * k + 2 * L + L + 2 * L);
            break;
        default: limbs = 3 * L + ks + 2 * L + 2 * L + 2 * (1 + L); break;
    }
    return limbs * c.n * 8;
}

u64 Server::submit(const OpPtr& op) {
    if (!op->ev || !op->a) fail(CKH_ERR_INVALID_PARAMS, "request needs an evaluator and an input");
    if (op->kind > CKH_OP_MUL_RELIN_RESCALE) fail(CKH_ERR_INVALID_PARAMS, "unknown op kind");
    if (op->kind == CKH_OP_MUL_RELIN_RESCALE && !op->b) fail(CKH_ERR_INVALID_PARAMS, "mul needs two inputs");
    if (op->t_arrival < 0) op->t_arrival = now();
    Context& c = op->ev->ctx;
    hip_check(hipSetDevice(c.device), "hipSetDevice");
    hip_check(hipEventCreateWithFlags(&op->ready, hipEventDisableTiming), "event");
    hip_check(hipEventRecord(op->ready, c.stream()), "event");
    std::lock_guard<std::mutex> lk(mu_);
    op->ticket = next_ticket_++;
    table_[op->ticket] = op;
    pending_.push_back(op);
    ++submitted_;
    if (first_arrival_ < 0 || op->t_arrival < first_arrival_) first_arrival_ = op->t_arrival;
    cv_.notify_all();
    return op->ticket;
}

void Server::fail_op(const OpPtr& op, ckh_status st, const std::string& msg) {
    op->status = st;
    op->err = msg;
    op->outs.clear();
}

// Runs one scheduler turn's admitted requests: grouped by evaluator, relinearizations (also the
// one inside mul+relin+rescale), hoisted rotations, the products and the rescales of
// mul+relin+rescale batched, on the next stream of the evaluator's context.
void Server::launch(std::vector<OpPtr>& ops) {
    std::vector<Evaluator*> evs;
    for (auto& op : ops)
        if (std::find(evs.begin(), evs.end(), op->ev.get()) == evs.end()) evs.push_back(op->ev.get());
    for (Evaluator* ev : evs) {
        Context& c = ev->ctx;
        hip_check(hipSetDevice(c.device), "hipSetDevice");
        c.set_service_stream(rr_[&c]++);
        hipStream_t s = c.stream();
        std::vector<OpPtr> mine;
        for (auto& op : ops)
            if (op->ev.get() == ev) {
                mine.push_back(op);
                (void)hipStreamWaitEvent(s, op->ready, 0);
            }
        // Injected failures and invalid requests fail alone; the rest run batched.
        std::vector<OpPtr> relin, mrr, hoisted;
        for (auto& op : mine) {
            if (op->inject_failure) {
                fail_op(op, CKH_ERR_DEVICE, "injected device failure");
                continue;
            }
            if (op->kind == CKH_OP_RELINEARIZE) relin.push_back(op);
            else if (op->kind == CKH_OP_MUL_RELIN_RESCALE) mrr.push_back(op);
            else hoisted.push_back(op);
        }
        // Hoisted rotations of every request in one batch (one decomposition launch sequence and
        // one switch program for all their steps); a failing member falls back to one call each.
        // Groups of at most c.hoisted_batch requests (tuning key "hoisted_batch"), so one
        // program's scratch (about 70 MB per step at N = 2^16 and 24 limbs) stays modest.
        const size_t per = std::max<u32>(c.hoisted_batch, 1);
        for (size_t g0 = 0; g0 < hoisted.size(); g0 += per) {
            const size_t g1 = std::min(hoisted.size(), g0 + per);
            std::vector<std::pair<const Ct*, std::vector<int32_t>>> items;
            for (size_t i = g0; i < g1; ++i) items.push_back({hoisted[i]->a.get(), hoisted[i]->steps});
            try {
                auto outs = ev->hoisted_rotate_batch(items);
                for (size_t i = g0; i < g1; ++i) hoisted[i]->outs = std::move(outs[i - g0]);
            } catch (const Error&) {
                for (size_t i = g0; i < g1; ++i) {
                    try {
                        hoisted[i]->outs = ev->hoisted_rotate(*hoisted[i]->a, hoisted[i]->steps);
                    } catch (const Error& e) {
                        fail_op(hoisted[i], e.status, e.msg);
                    }
                }
            }
        }
        // mul_no_relin first, every pair in one launch (a failing product: one call each, so it
        // fails its request alone).
        std::vector<CtPtr> prods;
        std::vector<OpPtr> mrr_ok;
        try {
            std::vector<std::pair<const Ct*, const Ct*>> pairs;
            for (auto& op : mrr) pairs.push_back({op->a.get(), op->b.get()});
            prods = ev->mul_no_relin_batch(pairs);
            mrr_ok = mrr;
        } catch (const Error&) {
            prods.clear();
            for (auto& op : mrr) {
                try {
                    prods.push_back(ev->mul_no_relin(*op->a, *op->b));
                    mrr_ok.push_back(op);
                } catch (const Error& e) {
                    fail_op(op, e.status, e.msg);
                }
            }
        }
        std::vector<const Ct*> batch;
        std::vector<OpPtr> owners;
        for (auto& op : relin) {
            batch.push_back(op->a.get());
            owners.push_back(op);
        }
        for (size_t i = 0; i < mrr_ok.size(); ++i) {
            batch.push_back(prods[i].get());
            owners.push_back(mrr_ok[i]);
        }
        std::vector<CtPtr> outs;
        try {
            outs = ev->relinearize_batch(batch);
        } catch (const Error&) {
            // One bad member: fall back to one call each so only that request fails.
            outs.assign(batch.size(), nullptr);
            for (size_t i = 0; i < batch.size(); ++i) {
                try {
                    outs[i] = ev->relinearize(*batch[i]);
                } catch (const Error& e2) {
                    fail_op(owners[i], e2.status, e2.msg);
                }
            }
        }
        // the rescales of the products, in one program (a failing one: one call each)
        std::vector<const Ct*> to_rescale;
        std::vector<size_t> idx;
        for (size_t i = 0; i < owners.size(); ++i) {
            if (!outs[i]) continue;
            if (owners[i]->kind == CKH_OP_RELINEARIZE) {
                owners[i]->outs = {outs[i]};
            } else {
                to_rescale.push_back(outs[i].get());
                idx.push_back(i);
            }
        }
        try {
            auto rs = ev->rescale_batch(to_rescale);
            for (size_t k = 0; k < idx.size(); ++k) owners[idx[k]]->outs = {rs[k]};
        } catch (const Error&) {
            for (size_t k = 0; k < idx.size(); ++k) {
                try {
                    owners[idx[k]]->outs = {ev->rescale(*to_rescale[k])};
                } catch (const Error& e) {
                    fail_op(owners[idx[k]], e.status, e.msg);
                }
            }
        }
        Batch b{&c, nullptr, mine};
        hip_check(hipEventCreateWithFlags(&b.done, hipEventDisableTiming), "event");
        if (hipEventRecord(b.done, s) != hipSuccess) {
            (void)hipEventDestroy(b.done);
            fail(CKH_ERR_DEVICE, "event record failed");
        }
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& op : mine) op->launched = true;
        inflight_.push_back(b);
        ++batches_;
    }
}

void Server::loop() {
    for (;;) {
        std::vector<OpPtr> admitted;
        {
            std::unique_lock<std::mutex> lk(mu_);
            // completions
            for (size_t i = 0; i < inflight_.size();) {
                Batch& b = inflight_[i];
                (void)hipSetDevice(b.ctx->device);
                const hipError_t q = hipEventQuery(b.done);
                if (q != hipErrorNotReady) {
                    // done, or the device reported an error: then the batch's requests fail
                    // instead of waiting for ever
                    const double t = now();
                    for (auto& op : b.ops) {
                        if (q != hipSuccess && op->status == CKH_OK)
                            fail_op(op, CKH_ERR_DEVICE, std::string("device error: ") + hipGetErrorString(q));
                        op->t_done = t;
                        op->done = true;
                        in_use_ -= op->footprint;
                        if (op->status == CKH_OK) {
                            ++completed_;
                            lat_ms_.push_back((t - op->t_arrival) * 1e3);
                        } else {
                            ++failed_;
                        }
                        last_done_ = std::max(last_done_, t);
                    }
                    (void)hipEventDestroy(b.done);
                    inflight_.erase(inflight_.begin() + i);
                    done_cv_.notify_all();
                } else {
                    ++i;
                }
            }
            if (stop_ && pending_.empty() && inflight_.empty()) return;
            // Admission in arrival order within the budget, while fewer than max_inflight batches
            // run: beyond that the device queue only grows (memory held, latency added, nothing
            // finished sooner), so arrivals wait here and join the next, larger batch.
            while (inflight_.size() < cfg_.max_inflight && !pending_.empty() && admitted.size() < cfg_.max_batch) {
                OpPtr op = pending_.front();
                if (op->t_arrival > now()) break;  // not arrived yet (pre-scheduled arrivals)
                u64 fp = 0;
                try {
                    fp = footprint(*op);
                } catch (...) {
                    fp = 0;
                }
                if (fp > cfg_.hbm_budget_bytes) {
                    pending_.pop_front();
                    fail_op(op, CKH_ERR_DEVICE, "request exceeds the HBM budget");
                    op->done = true;
                    op->t_done = now();
                    ++rejected_;
                    done_cv_.notify_all();
                    continue;
                }
                if (in_use_ + fp > cfg_.hbm_budget_bytes) break;
                pending_.pop_front();
                op->footprint = fp;
                in_use_ += fp;
                peak_ = std::max(peak_, in_use_);
                admitted.push_back(op);
            }
            if (admitted.empty()) {
                cv_.wait_for(lk, std::chrono::microseconds(inflight_.empty() ? 200 : 20));
                continue;
            }
        }
        try {
            const double t_l = now();
            launch(admitted);
            const double dt = (now() - t_l) * 1e3;
            std::lock_guard<std::mutex> lk(mu_);
            launch_ms_ += dt;
            launch_max_ms_ = std::max(launch_max_ms_, dt);
        } catch (const Error& e) {
            fail_admitted(admitted, e.status, e.msg);
        } catch (const std::exception& e) {  // host out of memory and the like: the loop keeps running
            fail_admitted(admitted, CKH_ERR_DEVICE, e.what());
        } catch (...) {
            fail_admitted(admitted, CKH_ERR_DEVICE, "unknown internal error");
        }
    }
}

// A turn whose launch threw: every request of it not yet completed fails with the error.
void Server::fail_admitted(std::vector<OpPtr>& admitted, ckh_status st, const std::string& msg) {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& op : admitted)
        if (!op->done && !op->launched) {  // launched ones complete with their batch
            fail_op(op, st, msg);
            op->done = true;
            in_use_ -= op->footprint;
            ++failed_;
        }
    done_cv_.notify_all();
}

bool Server::poll(u64 ticket, bool wait, OpPtr& op) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = table_.find(ticket);
    if (it == table_.end()) fail(CKH_ERR_INVALID_PARAMS, "unknown ticket");
    OpPtr o = it->second;
    if (wait) done_cv_.wait(lk, [&] { return o->done; });
    if (!o->done) return false;
    op = o;
    table_.erase(it);
    if (o->ready) (void)hipEventDestroy(o->ready);
    o->ready = nullptr;
    return true;
}

ckh_server_stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    ckh_server_stats s{};
    s.submitted = submitted_;
    s.completed = completed_;
    s.failed = failed_;
    s.rejected = rejected_;
    s.batches = batches_;
    s.peak_bytes = peak_;
    s.launch_ms = launch_ms_;
    s.launch_max_ms = launch_max_ms_;
    if (!lat_ms_.empty()) {
        std::vector<double> v = lat_ms_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) { return v[std::min(v.size() - 1, (size_t)std::ceil(p * v.size()) - 1)]; };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_ms = sum / v.size();
        s.max_ms = v.back();
        s.makespan_s = last_done_ - first_arrival_;
        s.ops_per_s = s.makespan_s > 0 ? completed_ / s.makespan_s : 0;
    }
    return s;
}

}  // namespace ckh
