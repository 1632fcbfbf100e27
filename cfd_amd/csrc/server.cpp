// Serving loop of the solver: requests (one mesh's system plus a run of time-step right-hand
// sides) are admitted in arrival order while their device footprint fits the memory budget,
// stay resident for all of their steps, and share every kernel launch with the other running
// solves through a SlotRunner (continuous batching).
//
// Host threads: the engine thread drives the device (one iteration boundary per step() of the
// slot runner, starting a request's next step in the slot its previous step freed); staging
// threads build admitted requests' systems on the device (upload, hierarchy, right-hand sides)
// and copy finished requests' results back, so that neither kind of host work stalls the
// launches, and several requests are staged at once. No counterpart in the reference, whose DeviceSolver serves one caller's solve
// at a time (src/gpu/src/device_solver.cpp:117-150).
#include "solver.hpp"

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

namespace cfdg {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

enum RequestState { RQ_WAITING, RQ_LOADING, RQ_READY, RQ_ACTIVE, RQ_RETIRING, RQ_DONE };

struct Request {
    unsigned long long id = 0;
    cfdg_server_request req{};
    std::vector<const double*> rhs;
    std::vector<double*> x_out;
    Clock::time_point arrival;
    Clock::time_point ready;    // the current step became ready
    Clock::time_point resident; // the system and right-hand sides became resident
    Clock::time_point finished;
    std::size_t reserved = 0;   // bytes counted against the budget
    std::unique_ptr<System> sys;
    DevBuf<double> rhs_dev;     // n_steps x n right-hand sides
    DevBuf<double> x_dev;       // n_steps x n solutions (when any is wanted)
    std::vector<double> tol;    // absolute tolerance of each step
    std::vector<cfdg_solve_stats> stats;
    std::vector<double> latency;
    int step = 0;
    Event staged;               // recorded after the request's last device work
    RequestState state = RQ_WAITING;
    cfdg_status status = CFDG_OK;
    std::string error;
};

cfdg_status status_of(const std::exception& e) {
    if (dynamic_cast<const InvalidArgument*>(&e) != nullptr) {
        return CFDG_ERR_INVALID_ARGUMENT;
    }
    if (dynamic_cast<const OutOfMemory*>(&e) != nullptr) {
        return CFDG_ERR_OUT_OF_MEMORY;
    }
    if (dynamic_cast<const CapacityError*>(&e) != nullptr) {
        return CFDG_ERR_CAPACITY;
    }
    if (dynamic_cast<const HipError*>(&e) != nullptr) {
        return CFDG_ERR_HIP;
    }
    return CFDG_ERR_INTERNAL;
}

} // namespace

class Server {
public:
    explicit Server(const cfdg_server_options& o) : opt_(o) {
        require(o.max_active_solves >= 1 && o.max_active_solves <= 4096,
                "server: max_active_solves must be in [1, 4096]");
        require(o.method == CFDG_METHOD_CG || o.method == CFDG_METHOD_BICGSTAB, "server: unknown method");
        require(o.precision == CFDG_PREC_FP64 || o.precision == CFDG_PREC_MIXED_FP32 ||
                    o.precision == CFDG_PREC_MIXED_BF16,
                "server: unknown precision");
        require(o.smooth_sweeps >= 0 && o.smooth_sweeps <= 64, "server: invalid smooth_sweeps");
        require(o.staging_threads >= 1 && o.staging_threads <= 64, "server: staging_threads must be in [1, 64]");
        require_device();
        budget_ = o.memory_budget_bytes;
        if (budget_ == 0) {
            std::size_t free_b = 0;
            std::size_t total_b = 0;
            hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
            budget_ = total_b / 10 * 9;
        }
        BatchOptions bo;
        bo.method = o.method;
        bo.use_amg = o.use_amg != 0 && o.method == CFDG_METHOD_CG;
        bo.precision = o.precision;
        amg_ = bo.use_amg;
        runner_ = std::make_unique<SlotRunner>(o.max_active_solves, bo, o.smooth_sweeps, stream_.get());
        owner_.assign(static_cast<std::size_t>(o.max_active_solves), nullptr);
        engine_ = std::thread([this] { engine_loop(); });
        for (int i = 0; i < o.staging_threads; ++i) {
            stagers_.emplace_back([this] { stager_loop(); });
        }
    }

    ~Server() {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return unfinished_ == 0; });
            stop_ = true;
        }
        cv_.notify_all();
        engine_.join();
        for (std::thread& t : stagers_) {
            t.join();
        }
    }

    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    unsigned long long submit(const cfdg_server_request& q) {
        require(q.n_steps >= 1 && q.rhs != nullptr, "server: a request needs at least one right-hand side");
        require(q.diag != nullptr, "server: null diagonal");
        require(q.rel_tolerance >= 0.0 && q.max_iterations >= 0, "server: invalid stopping criterion");
        for (int k = 0; k < q.n_steps; ++k) {
            require(q.rhs[k] != nullptr, "server: null right-hand side");
        }
        if (amg_) {
            require(q.amg != nullptr, "server: the server solves with AMG but the request has no hierarchy");
            require(q.amg->smooth_sweeps == opt_.smooth_sweeps,
                    "server: the hierarchy's smooth_sweeps differ from the server's");
        }
        auto r = std::make_unique<Request>();
        r->req = q;
        r->rhs.assign(q.rhs, q.rhs + q.n_steps);
        bool want_x = false;
        if (q.x_out != nullptr) {
            r->x_out.assign(q.x_out, q.x_out + q.n_steps);
            for (double* p : r->x_out) {
                want_x = want_x || p != nullptr;
            }
        }
        const std::size_t vec = static_cast<std::size_t>(q.pattern.n) * sizeof(double);
        r->reserved = estimate_footprint(q.pattern, amg_ ? q.amg : nullptr) +
                      vec * static_cast<std::size_t>(q.n_steps) * (want_x ? 2 : 1);
        r->stats.assign(static_cast<std::size_t>(q.n_steps), cfdg_solve_stats{});
        r->latency.assign(static_cast<std::size_t>(q.n_steps), 0.0);
        r->tol.assign(static_cast<std::size_t>(q.n_steps), 0.0);
        std::lock_guard<std::mutex> lk(mu_);
        if (!fatal_.empty()) {
            throw HipError("server stopped after an error: " + fatal_);
        }
        r->id = next_id_++;
        r->arrival = Clock::now();
        Request* p = r.get();
        reqs_[r->id] = std::move(r);
        waiting_.push_back(p);
        ++unfinished_;
        ++submitted_;
        cv_.notify_all();
        return p->id;
    }

    cfdg_status wait(unsigned long long id, double* latency_ms, std::string& err) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        require(it != reqs_.end(), "server: unknown request id (or already waited for)");
        Request* r = it->second.get();
        cv_.wait(lk, [&] { return r->state == RQ_DONE; });
        if (latency_ms != nullptr) {
            *latency_ms = ms_between(r->arrival, r->finished);
        }
        const cfdg_status st = r->status;
        err = r->error;
        if (fatal_.empty()) { // after a device error a thread may still hold it
            reqs_.erase(it);
        }
        return st;
    }

    cfdg_server_counters counters() {
        std::lock_guard<std::mutex> lk(mu_);
        cfdg_server_counters c{};
        c.submitted = submitted_;
        c.completed = completed_;
        c.solves = solves_;
        c.resident_bytes = resident_;
        c.peak_resident_bytes = peak_resident_;
        c.budget_bytes = budget_;
        c.active_solves = active_;
        c.peak_active_solves = peak_active_;
        c.waiting = static_cast<int>(waiting_.size());
        c.staging_ms = staging_ms_;
        c.retiring_ms = retiring_ms_;
        c.engine_busy_ms = engine_busy_ms_;
        return c;
    }

private:
    // ---- engine thread -------------------------------------------------------------------

    void begin_step(Request& r) {
        const std::size_t n = static_cast<std::size_t>(r.req.pattern.n);
        const std::size_t k = static_cast<std::size_t>(r.step);
        r.sys->copy_rhs_from(r.rhs_dev.get() + k * n, stream_.get());
        const cfdg_solve_control c{r.tol[k], r.req.max_iterations};
        const int slot = runner_->start(r.sys.get(), c);
        owner_[static_cast<std::size_t>(slot)] = &r;
    }

    // Hands a request to the staging thread (results back to the caller, memory freed).
    void retire(Request& r) {
        hip_check(hipEventRecord(r.staged.get(), stream_.get()), "event record");
        std::lock_guard<std::mutex> lk(mu_);
        r.finished = Clock::now();
        r.state = RQ_RETIRING;
        retiring_.push_back(&r);
        cv_.notify_all();
    }

    void on_done(const SlotRunner::Done& d) {
        Request& r = *owner_[static_cast<std::size_t>(d.slot)];
        owner_[static_cast<std::size_t>(d.slot)] = nullptr;
        const Clock::time_point now = Clock::now();
        const std::size_t k = static_cast<std::size_t>(r.step);
        r.stats[k] = d.stats;
        r.latency[k] = ms_between(r.ready, now);
        if (r.x_dev.size() > 0) {
            r.sys->copy_x_to(r.x_dev.get() + k * static_cast<std::size_t>(r.req.pattern.n), stream_.get());
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            ++solves_;
        }
        ++r.step;
        if (r.step < r.req.n_steps) {
            r.ready = now;
            begin_step(r);
        } else {
            retire(r);
        }
    }

    void engine_loop() {
        std::vector<SlotRunner::Done> done;
        try {
            for (;;) {
                std::vector<Request*> starting;
                {
                    std::unique_lock<std::mutex> lk(mu_);
                    if (!runner_->busy()) {
                        cv_.wait(lk, [&] { return stop_ || !ready_.empty(); });
                        if (ready_.empty()) {
                            return; // stop_: every request has completed
                        }
                    }
                    int free_slots = runner_->free_slots();
                    while (free_slots > 0 && !ready_.empty()) {
                        starting.push_back(ready_.front());
                        ready_.pop_front();
                        --free_slots;
                    }
                }
                for (Request* r : starting) {
                    r->state = RQ_ACTIVE;
                    r->ready = r->resident;
                    try {
                        begin_step(*r);
                    } catch (const InvalidArgument& e) {
                        r->status = CFDG_ERR_INVALID_ARGUMENT;
                        r->error = e.what();
                        retire(*r);
                    }
                }
                double busy = 0.0;
                if (runner_->busy()) {
                    const Clock::time_point t0 = Clock::now();
                    done.clear();
                    runner_->step(done);
                    for (const SlotRunner::Done& d : done) {
                        on_done(d);
                    }
                    busy = ms_between(t0, Clock::now());
                }
                std::lock_guard<std::mutex> lk(mu_);
                engine_busy_ms_ += busy;
                active_ = runner_->slots() - runner_->free_slots();
                peak_active_ = std::max(peak_active_, active_);
            }
        } catch (const std::exception& e) {
            fail_all(e.what());
        }
    }

    // ---- staging thread ------------------------------------------------------------------

    void load(Request& r) {
        const cfdg_server_request& q = r.req;
        auto sys = std::make_unique<System>(q.pattern);
        sys->assemble(q.diag, q.upper, q.lower);
        const bool bf16 = amg_ && opt_.precision == CFDG_PREC_MIXED_BF16;
        if (amg_) {
            sys->set_amg(*q.amg, false); // a request's few steps do not repay the fused products
            if (bf16) {
                sys->ensure_bf16();
            }
        }
        (void)sys->descriptor(amg_ && opt_.precision != CFDG_PREC_FP64, amg_, bf16); // allocates
        sys->set_x(q.x0);
        const std::size_t n = static_cast<std::size_t>(q.pattern.n);
        const std::size_t steps = static_cast<std::size_t>(q.n_steps);
        r.rhs_dev.resize(n * steps);
        for (std::size_t k = 0; k < steps; ++k) {
            hip_check(hipMemcpy(r.rhs_dev.get() + k * n, r.rhs[k], n * sizeof(double), hipMemcpyHostToDevice),
                      "H2D right-hand side");
            // |b| from 64 fixed chunks summed in order
            constexpr int chunks = 64;
            double part[chunks];
            const double* b = r.rhs[k];
            for (int c = 0; c < chunks; ++c) {
                double t = 0.0;
                for (std::size_t i = n * c / chunks; i < n * (c + 1) / chunks; ++i) {
                    t += b[i] * b[i];
                }
                part[c] = t;
            }
            double s = 0.0;
            for (int c = 0; c < chunks; ++c) {
                s += part[c];
            }
            r.tol[k] = q.rel_tolerance * std::sqrt(s);
        }
        bool want_x = false;
        for (double* p : r.x_out) {
            want_x = want_x || p != nullptr;
        }
        if (want_x) {
            r.x_dev.resize(n * steps);
        }
        hip_check(hipStreamSynchronize(sys->stream()), "staging sync");
        const std::size_t actual = sys->footprint() + r.rhs_dev.bytes() + r.x_dev.bytes();
        r.sys = std::move(sys);
        std::lock_guard<std::mutex> lk(mu_);
        resident_ = resident_ - r.reserved + actual;
        peak_resident_ = std::max(peak_resident_, resident_);
        r.reserved = actual;
        if (!fatal_.empty()) {
            return; // already failed; its memory goes with the server
        }
        r.state = RQ_READY;
        r.resident = Clock::now();
        ready_.push_back(&r);
        cv_.notify_all();
    }

    void finish(Request& r) {
        hip_check(hipEventSynchronize(r.staged.get()), "event sync");
        const std::size_t n = static_cast<std::size_t>(r.req.pattern.n);
        if (r.status == CFDG_OK) {
            for (std::size_t k = 0; k < r.x_out.size(); ++k) {
                if (r.x_out[k] != nullptr) {
                    hip_check(hipMemcpy(r.x_out[k], r.x_dev.get() + k * n, n * sizeof(double),
                                        hipMemcpyDeviceToHost),
                              "D2H solution");
                }
            }
            for (std::size_t k = 0; k < r.stats.size(); ++k) {
                if (r.req.stats_out != nullptr) {
                    r.req.stats_out[k] = r.stats[k];
                }
                if (r.req.step_latency_ms != nullptr) {
                    r.req.step_latency_ms[k] = r.latency[k];
                }
            }
        }
    }

    void release(Request& r) {
        r.sys.reset();
        r.rhs_dev.resize(0);
        r.x_dev.resize(0);
        std::lock_guard<std::mutex> lk(mu_);
        resident_ -= r.reserved;
        r.reserved = 0;
        if (r.state != RQ_DONE) {
            r.state = RQ_DONE;
            --unfinished_;
            ++completed_;
        }
        cv_.notify_all();
    }

    void stager_loop() {
        set_host_parallel(false); // several staging threads run at once
        for (;;) {
            Request* job = nullptr;
            bool finishing = false;
            {
                std::unique_lock<std::mutex> lk(mu_);
                for (;;) {
                    if (!retiring_.empty()) {
                        job = retiring_.front();
                        retiring_.pop_front();
                        finishing = true;
                        break;
                    }
                    if (!waiting_.empty() && fatal_.empty()) {
                        Request* front = waiting_.front();
                        if (front->reserved > budget_) { // can never fit
                            waiting_.pop_front();
                            front->status = CFDG_ERR_CAPACITY;
                            front->error = "server: request footprint exceeds the memory budget";
                            front->finished = Clock::now();
                            front->state = RQ_DONE;
                            --unfinished_;
                            ++completed_;
                            cv_.notify_all();
                            continue;
                        }
                        if (resident_ + front->reserved <= budget_) {
                            waiting_.pop_front();
                            job = front;
                            job->state = RQ_LOADING;
                            resident_ += job->reserved;
                            peak_resident_ = std::max(peak_resident_, resident_);
                            break;
                        }
                    }
                    if (stop_) {
                        return;
                    }
                    cv_.wait(lk);
                }
            }
            const Clock::time_point t0 = Clock::now();
            try {
                if (finishing) {
                    finish(*job);
                } else {
                    load(*job);
                    std::lock_guard<std::mutex> lk(mu_);
                    staging_ms_ += ms_between(t0, Clock::now());
                    continue;
                }
            } catch (const std::exception& e) {
                job->status = status_of(e);
                job->error = e.what();
                job->finished = Clock::now();
            }
            release(*job);
            std::lock_guard<std::mutex> lk(mu_);
            (finishing ? retiring_ms_ : staging_ms_) += ms_between(t0, Clock::now());
        }
    }

    // A device error ends the server: every request still open fails with it.
    void fail_all(const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu_);
        fatal_ = msg;
        for (auto& kv : reqs_) {
            Request& r = *kv.second;
            if (r.state != RQ_DONE) {
                r.status = CFDG_ERR_HIP;
                r.error = "server stopped after an error: " + msg;
                r.finished = Clock::now();
                r.state = RQ_DONE;
                --unfinished_;
            }
        }
        waiting_.clear();
        ready_.clear();
        retiring_.clear();
        cv_.notify_all();
    }

    cfdg_server_options opt_;
    bool amg_ = true;
    std::size_t budget_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::map<unsigned long long, std::unique_ptr<Request>> reqs_;
    std::deque<Request*> waiting_;
    std::deque<Request*> ready_;
    std::deque<Request*> retiring_;
    unsigned long long next_id_ = 1;
    unsigned long long submitted_ = 0;
    unsigned long long completed_ = 0;
    unsigned long long solves_ = 0;
    long long unfinished_ = 0;
    std::size_t resident_ = 0;
    std::size_t peak_resident_ = 0;
    int active_ = 0;
    int peak_active_ = 0;
    double staging_ms_ = 0.0;
    double retiring_ms_ = 0.0;
    double engine_busy_ms_ = 0.0;
    bool stop_ = false;
    std::string fatal_;
    // Declared after the requests so that the runner (which drains its stream) goes first.
    Stream stream_;
    std::unique_ptr<SlotRunner> runner_;
    std::vector<Request*> owner_; // per slot, engine thread only
    std::thread engine_;
    std::vector<std::thread> stagers_;
};

} // namespace cfdg

using namespace cfdg;

struct cfdg_server {
    std::unique_ptr<Server> s;
};

extern "C" cfdg_status cfdg_estimate_footprint(const cfdg_ldu_pattern* pattern, const cfdg_amg_view* amg,
                                               size_t* bytes) {
    return guarded([&] {
        require(pattern != nullptr && bytes != nullptr, "null argument");
        *bytes = estimate_footprint(*pattern, amg);
    });
}

extern "C" cfdg_status cfdg_server_default_options(cfdg_server_options* out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = cfdg_server_options{};
        out->max_active_solves = 64;
        out->method = CFDG_METHOD_CG;
        out->use_amg = 1;
        out->precision = CFDG_PREC_MIXED_FP32;
        out->smooth_sweeps = 2;
        out->memory_budget_bytes = 0;
        out->staging_threads = 4;
    });
}

extern "C" cfdg_status cfdg_server_create(const cfdg_server_options* options, cfdg_server** out) {
    return guarded([&] {
        require(options != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        auto h = std::make_unique<cfdg_server>();
        h->s = std::make_unique<Server>(*options);
        *out = h.release();
    });
}

extern "C" cfdg_status cfdg_server_submit(cfdg_server* server, const cfdg_server_request* request,
                                          unsigned long long* id) {
    return guarded([&] {
        require(server != nullptr && request != nullptr && id != nullptr, "null argument");
        *id = server->s->submit(*request);
    });
}

extern "C" cfdg_status cfdg_server_wait(cfdg_server* server, unsigned long long id, double* latency_ms) {
    cfdg_status st = CFDG_OK;
    std::string err;
    const cfdg_status call = guarded([&] {
        require(server != nullptr, "null argument");
        st = server->s->wait(id, latency_ms, err);
    });
    if (call != CFDG_OK) {
        return call;
    }
    if (st != CFDG_OK) {
        set_last_error(err);
    }
    return st;
}

extern "C" cfdg_status cfdg_server_query(cfdg_server* server, cfdg_server_counters* out) {
    return guarded([&] {
        require(server != nullptr && out != nullptr, "null argument");
        *out = server->s->counters();
    });
}

extern "C" cfdg_status cfdg_server_destroy(cfdg_server* server) {
    return guarded([&] { delete server; });
}
