// server.cpp -- continuous batching of MAX-CUT solve requests.
//
// A worker thread owns the device loop and keeps two slices queued on the
// context's stream: every admitted chain advances by `slice` sweeps per slice,
// in one persistent launch per kernel variant.  While the older slice runs,
// the loop admits queued requests whose HBM footprint fits the budget, in
// arrival order with a bounded bypass of requests that do not fit yet (CSR
// upload, chain-state allocation and the initial-word kernel on the side
// stream, all from a stream-ordered pool, so nothing waits for the running
// kernels), and once it has finished, retires the requests whose last sweeps
// it held (results copied to the caller's buffers on the side stream, device
// state returned to the pool) while the newer slice keeps the device busy.
// Which requests end in a slice is known when it is enqueued: the host
// mirrors every chain's step count.  A chain's arithmetic does not depend on
// when it was admitted or how its sweeps were sliced, so every result equals
// the one-shot result for the same inputs.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <utility>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../../include/rfmf.h"
#include "../host/host_common.hpp"
#include "capi_internal.hpp"
#include "engine.hpp"

using rfmf::Error;
using rfmf::Status;

namespace {

struct Request {
    int64_t id = 0;
    rfmf::HostGraph graph;   // prepared on submit, moved to the device at admission
    rfmf::HostJob host;      // prepared on submit (validation, gate, initial cuts)
    int64_t n = 0, m2 = 0;
    std::vector<int8_t> spins;
    std::vector<uint64_t> seeds;
    std::vector<int64_t> runEnd;
    std::vector<int32_t> runVal;
    std::vector<int64_t> runOff;
    rfmf::JobSpec spec;
    int64_t footprint = 0;
    double tSubmit = -1, tAdmit = -1, tDone = -1;
    int bypassed = 0;         // later requests admitted while this one waited
    bool done = false;
    bool failed = false;
    std::string error;
    int32_t graphId = 0;
    std::unique_ptr<rfmf::ResidentJob> job;
};

}  // namespace

struct rfmf_server {
    rfmf::Context* ctx = nullptr;
    int64_t budget = 0;
    int64_t slice = 0;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::mutex mu;
    std::condition_variable wake, doneCv;
    std::deque<std::shared_ptr<Request>> queue;
    std::vector<std::shared_ptr<Request>> active;
    std::unordered_map<int64_t, std::shared_ptr<Request>> all;
    int64_t nextId = 1;
    int64_t resident = 0;
    bool stopping = false;
    rfmf_server_stats stats{};
    std::thread worker;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // A queued request that does not fit yet can be passed by at most this
    // many later requests that do; then nothing passes it until it fits.
    static constexpr int kMaxBypass = 8;

    // Pops the queued requests that fit the remaining budget, in arrival
    // order.  A request that does not fit stays queued and the scan goes on
    // behind it (small requests need not wait for a large one to fit), but
    // each admission past it counts against its bypass allowance; a request
    // whose allowance is spent blocks everything behind it, so it cannot
    // starve.
    std::vector<std::shared_ptr<Request>> takeAdmissible() {
        std::vector<std::shared_ptr<Request>> out;
        std::vector<Request*> waiting;   // requests ahead of the scan that did not fit
        for (auto it = queue.begin(); it != queue.end();) {
            // Nothing passes a waiting request whose allowance is spent.
            bool spent = false;
            for (const Request* w : waiting) spent = spent || w->bypassed >= kMaxBypass;
            if (spent) break;
            Request& r = **it;
            if (resident + r.footprint <= budget) {
                resident += r.footprint;
                for (Request* w : waiting) w->bypassed += 1;
                out.push_back(*it);
                it = queue.erase(it);
            } else {
                waiting.push_back(&r);
                ++it;
            }
        }
        return out;
    }

    void admitOne(const std::shared_ptr<Request>& r) {
        try {
            std::lock_guard<std::mutex> lk(ctx->mutex());
            r->graphId = ctx->uploadGraph(std::move(r->graph), ctx->sideStream());
            r->spec.graph = ctx->graph(r->graphId);
            r->job = ctx->placeJob(r->host, r->spec.graph, ctx->sideStream(), /*streamOrdered=*/true);
        } catch (const std::exception& e) {
            r->failed = true;
            r->error = e.what();
        }
        std::lock_guard<std::mutex> lk(mu);
        r->tAdmit = now();
        if (r->failed) {
            finish(r);
        } else {
            active.push_back(r);
        }
    }

    // Called with mu held, once no queued slice reads the request's buffers.
    void finish(const std::shared_ptr<Request>& r) {
        r->tDone = now();
        r->done = true;
        {
            std::lock_guard<std::mutex> lk(ctx->mutex());
            r->job.reset();   // stream-ordered frees on the side stream
            if (r->graphId) {
                try {
                    ctx->releaseGraph(r->graphId, /*wait=*/false);
                } catch (...) {
                }
                r->graphId = 0;
            }
        }
        resident -= r->footprint;
        stats.completed += 1;
        doneCv.notify_all();
    }

    struct InFlight {
        rfmf::Context::SliceHandle h;
        std::vector<std::shared_ptr<Request>> finishing;   // requests whose last sweeps it holds
    };

    static bool allEnqueued(const Request& r) {
        for (uint64_t s : r.job->stepsDone)
            if (s < r.job->totalSteps) return false;
        return true;
    }

    // Enqueues one slice over every active request; requests that reach
    // their last sweep in it leave `active`.
    void enqueueSlice(std::deque<InFlight>& inflight) {
        std::vector<rfmf::ResidentJob*> jobs;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto& r : active) jobs.push_back(r->job.get());
        }
        InFlight f;
        rfmf::LaunchKnobs knobs;
        knobs.sliceSweeps = slice;
        {
            std::lock_guard<std::mutex> lk(ctx->mutex());
            f.h = ctx->launchSlice(jobs, knobs);
        }
        std::lock_guard<std::mutex> lk(mu);
        std::vector<std::shared_ptr<Request>> still;
        for (auto& r : active) (allEnqueued(*r) ? f.finishing : still).push_back(r);
        active.swap(still);
        inflight.push_back(std::move(f));
    }

    // Waits for the oldest queued slice and hands its finished requests back.
    void retireOldest(std::deque<InFlight>& inflight) {
        InFlight f = std::move(inflight.front());
        inflight.pop_front();
        float ms = 0.0f;
        {
            std::lock_guard<std::mutex> lk(ctx->mutex());
            ms = ctx->waitSlice(f.h);
        }
        for (auto& r : f.finishing) {
            try {
                std::lock_guard<std::mutex> lk(ctx->mutex());
                ctx->collect(*r->job, r->spec, ctx->sideStream());
            } catch (const std::exception& e) {
                r->failed = true;
                r->error = e.what();
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        if (!f.h.empty) stats.slices += 1;
        stats.busy_kernel_ms += ms;
        for (auto& r : f.finishing) {
            stats.chain_steps += static_cast<double>(r->job->totalSteps) * static_cast<double>(r->job->chains);
            finish(r);
        }
    }

    void loop() {
        std::deque<InFlight> inflight;
        for (;;) {
            bool haveActive = false;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (inflight.empty())
                    wake.wait(lk, [&] { return stopping || !queue.empty() || !active.empty(); });
                if (stopping && queue.empty() && active.empty() && inflight.empty()) return;
                haveActive = !active.empty();
            }
            // Keep two slices queued so the device never waits for the host.
            if (haveActive && inflight.size() < 2) enqueueSlice(inflight);
            // Admit while the queued slices run.
            std::vector<std::shared_ptr<Request>> admitNow;
            {
                std::lock_guard<std::mutex> lk(mu);
                admitNow = takeAdmissible();
            }
            for (auto& r : admitNow) admitOne(r);
            {
                std::lock_guard<std::mutex> lk(mu);
                haveActive = !active.empty();
            }
            if (haveActive && inflight.size() < 2) continue;   // room for another slice
            if (!inflight.empty()) retireOldest(inflight);
        }
    }
};

extern "C" {

rfmf_status rfmf_server_create(rfmf_ctx* ctx, int64_t hbm_budget_bytes, int64_t slice_sweeps,
                               rfmf_server** out) {
    if (!ctx || !out || hbm_budget_bytes <= 0 || slice_sweeps <= 0) return RFMF_ERR_INVALID_ARG;
    auto s = std::make_unique<rfmf_server>();
    s->ctx = ctx->ctx.get();
    s->budget = hbm_budget_bytes;
    s->slice = slice_sweeps;
    rfmf_server* raw = s.get();
    s->worker = std::thread([raw] { raw->loop(); });
    *out = s.release();
    return RFMF_OK;
}

rfmf_status rfmf_server_submit(rfmf_server* srv, const rfmf_csr* g, const rfmf_job* job, int64_t* request_id) {
    if (!srv || !g || !job || !request_id || !g->offsets || g->n <= 0 || job->chains <= 0 ||
        !job->initial_spins || !job->solver_seeds || !job->beta.run_end || !job->beta.run_val ||
        !job->beta.chain_run_off || job->beta.n_runs < 1 || job->n_phases < 0 ||
        (job->n_phases > 0 && (!job->phases || !job->iters)))
        return RFMF_ERR_INVALID_ARG;
    auto r = std::make_shared<Request>();
    r->n = g->n;
    r->m2 = g->m2;
    r->spins.assign(job->initial_spins, job->initial_spins + job->chains * g->n);
    r->seeds.assign(job->solver_seeds, job->solver_seeds + job->chains);
    r->runEnd.assign(job->beta.run_end, job->beta.run_end + job->beta.n_runs);
    r->runVal.assign(job->beta.run_val, job->beta.run_val + job->beta.n_runs);
    r->runOff.assign(job->beta.chain_run_off, job->beta.chain_run_off + job->chains);
    rfmf::JobSpec& s = r->spec;
    s.chains = job->chains;
    s.initialSpins = r->spins.data();
    s.seeds = r->seeds.data();
    s.runEnd = r->runEnd.data();
    s.runVal = r->runVal.data();
    s.chainRunOff = r->runOff.data();
    s.nRuns = job->beta.n_runs;
    s.phases.assign(job->phases, job->phases + job->n_phases);
    s.iters.assign(job->iters, job->iters + job->n_phases);
    s.endSpins = job->end_spins;
    s.bestSpins = job->best_spins;
    s.bestVals = job->best_vals;
    s.finalVals = job->final_vals;
    const int bits = rfmf::wordBitsFor(rfmf_max_weighted_degree(g));
    r->footprint = rfmf::footprintBytes(g->n, g->m2, job->chains, job->beta.n_runs, job->n_phases,
                                        bits > 32 ? bits : 32);
    auto reject = [&](rfmf_status st, const std::string& why) {
        rfmf::setLastError(why);
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stats.submitted += 1;
        srv->stats.rejected += 1;
        return st;
    };
    if (r->footprint > srv->budget)
        return reject(RFMF_ERR_ADMISSION, "request footprint " + std::to_string(r->footprint) +
                                              " B exceeds the server's HBM budget");
    // The host half of admission runs here, on the caller's thread (CSR and
    // job validation, the int32 gate, initial cuts): the device loop then
    // only allocates, uploads and launches.
    try {
        r->graph = rfmf::prepareGraph(g->offsets, g->edge_to, g->edge_weight, g->n, g->m2);
        r->host = srv->ctx->prepareJob(s, r->graph);
    } catch (const rfmf::Error& e) {
        return reject(static_cast<rfmf_status>(e.status()), e.what());
    } catch (const std::exception& e) {
        return reject(RFMF_ERR_INVALID_ARG, e.what());
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stopping) return RFMF_ERR_STATE;
    srv->stats.submitted += 1;
    r->id = srv->nextId++;
    r->tSubmit = srv->now();
    srv->all.emplace(r->id, r);
    srv->queue.push_back(r);
    *request_id = r->id;
    srv->wake.notify_all();
    return RFMF_OK;
}

rfmf_status rfmf_server_poll(rfmf_server* srv, int64_t request_id, int32_t* done, double* t_submit,
                             double* t_admit, double* t_done) {
    if (!srv) return RFMF_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    const auto it = srv->all.find(request_id);
    if (it == srv->all.end()) return RFMF_ERR_NOT_FOUND;
    const Request& r = *it->second;
    if (done) *done = r.done ? 1 : 0;
    if (t_submit) *t_submit = r.tSubmit;
    if (t_admit) *t_admit = r.tAdmit;
    if (t_done) *t_done = r.tDone;
    if (r.done && r.failed) {
        rfmf::setLastError(r.error);
        return RFMF_ERR_STATE;
    }
    return RFMF_OK;
}

rfmf_status rfmf_server_wait(rfmf_server* srv, int64_t request_id, double timeout_s) {
    if (!srv) return RFMF_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(srv->mu);
    const auto it = srv->all.find(request_id);
    if (it == srv->all.end()) return RFMF_ERR_NOT_FOUND;
    std::shared_ptr<Request> r = it->second;
    const bool ok = srv->doneCv.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; });
    if (!ok) return RFMF_ERR_STATE;
    return r->failed ? RFMF_ERR_STATE : RFMF_OK;
}

rfmf_status rfmf_server_stats_get(rfmf_server* srv, rfmf_server_stats* out) {
    if (!srv || !out) return RFMF_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    int64_t chains = 0;
    for (const auto& r : srv->active) chains += r->spec.chains;
    out->active_chains = chains;
    out->resident_bytes = srv->resident;
    return RFMF_OK;
}

void rfmf_server_destroy(rfmf_server* srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
        srv->wake.notify_all();
    }
    if (srv->worker.joinable()) srv->worker.join();
    delete srv;
}

}  // extern "C"
