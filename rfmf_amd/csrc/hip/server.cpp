// server.cpp -- continuous batching of MAX-CUT solve requests.
//
// A worker thread owns the device loop.  Each iteration it enqueues one slice
// (every admitted chain advances by `slice` sweeps in one persistent launch
// per kernel variant), and while that slice runs it admits queued requests
// whose HBM footprint fits the budget: the CSR upload, the chain-state
// allocation and the initial-word kernel go to the side stream and overlap the
// running slice.  After the slice it retires finished requests (results
// copied to the caller's buffers, device state freed) and loops.  A chain's
// arithmetic does not depend on when it was admitted or how its sweeps were
// sliced, so every result equals the one-shot result for the same inputs.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
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
    std::vector<uint32_t> off;
    std::vector<int32_t> to, w;
    int64_t n = 0, m2 = 0;
    std::vector<int8_t> spins;
    std::vector<uint64_t> seeds;
    std::vector<int64_t> runEnd;
    std::vector<int32_t> runVal;
    std::vector<int64_t> runOff;
    rfmf::JobSpec spec;
    int64_t footprint = 0;
    double tSubmit = -1, tAdmit = -1, tDone = -1;
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

    // Pops every queued request that fits the remaining budget (FIFO; a
    // request that does not fit yet blocks the ones behind it, so arrival
    // order is kept).
    std::vector<std::shared_ptr<Request>> takeAdmissible() {
        std::vector<std::shared_ptr<Request>> out;
        while (!queue.empty() && resident + queue.front()->footprint <= budget) {
            resident += queue.front()->footprint;
            out.push_back(queue.front());
            queue.pop_front();
        }
        return out;
    }

    void admitOne(const std::shared_ptr<Request>& r) {
        try {
            std::lock_guard<std::mutex> lk(ctx->mutex());
            r->graphId = ctx->uploadGraph(r->off.data(), r->to.data(), r->w.data(), r->n, r->m2);
            r->spec.graph = ctx->graph(r->graphId);
            r->job = ctx->admit(r->spec, ctx->sideStream());
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

    // Called with mu held.
    void finish(const std::shared_ptr<Request>& r) {
        r->tDone = now();
        r->done = true;
        r->job.reset();
        if (r->graphId) {
            try {
                std::lock_guard<std::mutex> lk(ctx->mutex());
                ctx->releaseGraph(r->graphId);
            } catch (...) {
            }
            r->graphId = 0;
        }
        resident -= r->footprint;
        stats.completed += 1;
        doneCv.notify_all();
    }

    void loop() {
        for (;;) {
            std::vector<std::shared_ptr<Request>> admitNow;
            std::vector<rfmf::ResidentJob*> jobs;
            {
                std::unique_lock<std::mutex> lk(mu);
                wake.wait(lk, [&] { return stopping || !queue.empty() || !active.empty(); });
                if (stopping && queue.empty() && active.empty()) return;
                if (active.empty()) {
                    admitNow = takeAdmissible();
                } else {
                    for (auto& r : active) jobs.push_back(r->job.get());
                }
            }
            if (jobs.empty()) {
                for (auto& r : admitNow) admitOne(r);
                continue;
            }
            rfmf::Context::SliceHandle h;
            rfmf::LaunchKnobs knobs;
            knobs.sliceSweeps = slice;
            {
                std::lock_guard<std::mutex> lk(ctx->mutex());
                h = ctx->launchSlice(jobs, knobs);
            }
            // Admit while the slice runs.
            {
                std::lock_guard<std::mutex> lk(mu);
                admitNow = takeAdmissible();
            }
            for (auto& r : admitNow) admitOne(r);
            float ms = 0.0f;
            {
                std::lock_guard<std::mutex> lk(ctx->mutex());
                ms = ctx->waitSlice(h);
            }
            std::lock_guard<std::mutex> lk(mu);
            stats.slices += 1;
            stats.busy_kernel_ms += ms;
            std::vector<std::shared_ptr<Request>> still;
            for (auto& r : active) {
                bool fin = true;
                for (uint64_t s : r->job->stepsDone) fin &= s >= r->job->totalSteps;
                if (!fin) { still.push_back(r); continue; }
                stats.chain_steps += static_cast<double>(r->job->totalSteps) * static_cast<double>(r->job->chains);
                try {
                    std::lock_guard<std::mutex> lk2(ctx->mutex());
                    ctx->collect(*r->job, r->spec, ctx->stream());
                } catch (const std::exception& e) {
                    r->failed = true;
                    r->error = e.what();
                }
                finish(r);
            }
            active.swap(still);
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
    r->off.assign(g->offsets, g->offsets + g->n + 1);
    r->to.assign(g->edge_to, g->edge_to + g->m2);
    r->w.assign(g->edge_weight, g->edge_weight + g->m2);
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
    r->footprint = rfmf_footprint_bytes(g->n, g->m2, job->chains, job->beta.n_runs, job->n_phases);
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stats.submitted += 1;
    if (r->footprint > srv->budget) {
        srv->stats.rejected += 1;
        return RFMF_ERR_ADMISSION;
    }
    if (srv->stopping) return RFMF_ERR_STATE;
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
