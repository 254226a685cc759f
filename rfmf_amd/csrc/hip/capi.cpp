// capi.cpp -- the C ABI (include/rfmf.h): status mapping, host helpers,
// the one-shot annealGpu replacement, resident contexts and batches.
#include "../../../include/rfmf.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/host_common.hpp"
#include "../host/immutable_graph.hpp"
#include "../host/max_cut.hpp"
#include "../host/read_gset.hpp"
#include "capi_internal.hpp"
#include "engine.hpp"

using rfmf::Error;
using rfmf::Status;

using rfmf::guarded;

namespace {

thread_local std::string g_lastError;

void requireCsr(const rfmf_csr* g) {
    if (!g || !g->offsets || g->n < 0 || g->m2 < 0 || (g->m2 > 0 && (!g->edge_to || !g->edge_weight)))
        throw Error(Status::InvalidArg, "invalid CSR view");
}

}  // namespace

struct rfmf_graph {
    anneallib::AnnealCsr csr;
};

namespace rfmf {
void setLastError(const std::string& msg) { g_lastError = msg; }
}  // namespace rfmf

int64_t rfmf::footprintBytes(int64_t n, int64_t m2, int64_t chains, int64_t beta_runs, int64_t n_phases,
                             int bits) {
    return (n + 1) * 4 + m2 * 8 + chains * (rfmf::dev::wordBytes(n, bits) + rfmf::dev::bestBytes(n) + 32) +
           beta_runs * 12 + n_phases * 12 + chains * 8;
}

namespace rfmf {

void passThrough(const rfmf_csr* g, const int8_t* init, int64_t C, int8_t* end, int8_t* best, float* ms,
                 int64_t* bestVals, int64_t* finalVals) {
    const int64_t n = g->n;
    if (end) std::memcpy(end, init, static_cast<size_t>(C * n));
    if (best) std::memcpy(best, init, static_cast<size_t>(C * n));
    if (ms) *ms = 0.0f;
    for (int64_t c = 0; c < C; ++c) {
        const int64_t v = anneallib::solver::cutValueRaw(g->offsets, g->edge_to, g->edge_weight, n, init + c * n);
        if (bestVals) bestVals[c] = v;
        if (finalVals) finalVals[c] = v;
    }
}

void checkOneShot(const rfmf_csr* g, const int8_t* init, int64_t C, const uint64_t* phases, const uint64_t* iters,
                  int64_t nPhases, const uint64_t* seeds) {
    requireCsr(g);
    if (C <= 0 || !seeds || (g->n > 0 && !init))
        throw Error(Status::InvalidArg, "annealGpu needs non-empty, equal-length t0s, solver seeds and "
                                        "initial spins");
    if (nPhases < 0 || (nPhases > 0 && (!phases || !iters)))
        throw Error(Status::InvalidArg, "invalid phases / iters");
}

}  // namespace rfmf

extern "C" {

const char* rfmf_status_string(rfmf_status s) {
    switch (s) {
        case RFMF_OK: return "ok";
        case RFMF_ERR_INVALID_ARG: return "invalid argument";
        case RFMF_ERR_UNSUPPORTED: return "unsupported";
        case RFMF_ERR_HIP: return "HIP runtime error";
        case RFMF_ERR_NO_DEVICE: return "no device";
        case RFMF_ERR_OUT_OF_MEMORY: return "out of memory";
        case RFMF_ERR_BETA_CAP: return "beta run cap exceeded";
        case RFMF_ERR_STEP_CAP: return "total sweep cap exceeded";
        case RFMF_ERR_IO: return "I/O error";
        case RFMF_ERR_NOT_FOUND: return "not found";
        case RFMF_ERR_ADMISSION: return "admission rejected";
        case RFMF_ERR_STATE: return "invalid state";
    }
    return "unknown status";
}

const char* rfmf_last_error(void) { return g_lastError.c_str(); }

const char* rfmf_version(void) { return "rfmf 0.1.0 gfx950"; }

// ---------------------------------------------------------------- graphs

rfmf_status rfmf_graph_read_gset(const char* path, double forced_scale, rfmf_graph** out,
                                 rfmf_read_stats* stats) {
    return guarded([&] {
        if (!path || !out) throw Error(Status::InvalidArg, "null path or output");
        anneallib::Graph g;
        anneallib::ReadStats rs;
        if (!anneallib::read_instance_from_file(path, g, &rs, forced_scale))
            throw Error(Status::Io, std::string("failed to read graph: ") + path);
        removeDegreeZeroVertices(g);
        auto h = std::make_unique<rfmf_graph>();
        h->csr = anneallib::AnnealCsr(g);
        if (stats) {
            stats->num_vertices = rs.numVertices;
            stats->num_edges = rs.numEdges;
            stats->zero_dropped = rs.zeroDropped;
            stats->float_weights = rs.floatWeights ? 1 : 0;
            stats->pad = 0;
            stats->scale = rs.scale;
            stats->max_abs_weighted_degree = rs.maxAbsWeightedDegree;
        }
        *out = h.release();
    });
}

rfmf_status rfmf_graph_from_edges(const int64_t* u, const int64_t* v, const int64_t* w, int64_t m,
                                  rfmf_graph** out) {
    return guarded([&] {
        if (!out || m < 0 || (m > 0 && (!u || !v || !w))) throw Error(Status::InvalidArg, "invalid edge list");
        anneallib::Graph g;
        for (int64_t i = 0; i < m; ++i) {
            if (w[i] < INT32_MIN || w[i] > INT32_MAX)
                throw Error(Status::InvalidArg, "edge weight outside int32");
            g.addEdge(u[i], v[i], w[i]);
        }
        removeDegreeZeroVertices(g);
        auto h = std::make_unique<rfmf_graph>();
        h->csr = anneallib::AnnealCsr(g);
        *out = h.release();
    });
}

void rfmf_graph_free(rfmf_graph* g) { delete g; }

rfmf_status rfmf_graph_view(const rfmf_graph* g, rfmf_csr* view) {
    return guarded([&] {
        if (!g || !view) throw Error(Status::InvalidArg, "null graph or view");
        view->offsets = g->csr.offsets().data();
        view->edge_to = g->csr.edgeTo().data();
        view->edge_weight = g->csr.edgeWeight().data();
        view->n = static_cast<int64_t>(g->csr.numVertices());
        view->m2 = static_cast<int64_t>(g->csr.edgeTo().size());
    });
}

const int64_t* rfmf_graph_original_ids(const rfmf_graph* g) {
    return g ? g->csr.newToOld().data() : nullptr;
}

int64_t rfmf_cut_value(const rfmf_csr* g, const int8_t* spin) {
    if (!g || !spin || !g->offsets) return 0;
    return anneallib::solver::cutValueRaw(g->offsets, g->edge_to, g->edge_weight, g->n, spin);
}

int64_t rfmf_max_weighted_degree(const rfmf_csr* g) {
    if (!g || !g->offsets) return 0;
    int64_t B = 0;
    for (int64_t v = 0; v < g->n; ++v) {
        int64_t s = 0;
        for (uint32_t i = g->offsets[v]; i < g->offsets[v + 1]; ++i)
            s += std::llabs(static_cast<long long>(g->edge_weight[i]));
        B = std::max(B, s);
    }
    return B;
}

int32_t rfmf_score_safe_int32(int64_t total_steps, double t0, int64_t delta_bound) {
    return annealScoreSafeInt32(total_steps, t0, delta_bound) ? 1 : 0;
}

int32_t rfmf_beta_fixed_for_step(uint64_t step, double t0) { return betaFixedForStep(step, t0); }

rfmf_status rfmf_total_steps(const uint64_t* iters, int64_t n, int64_t* total) {
    return guarded([&] {
        if (!total || n < 0 || (n > 0 && !iters)) throw Error(Status::InvalidArg, "invalid iters");
        *total = annealTotalSteps(std::vector<uint64_t>(iters, iters + n));
    });
}

rfmf_status rfmf_initial_spins(uint64_t raw_seed, int64_t n, int8_t* out) {
    return guarded([&] {
        if (n < 0 || (n > 0 && !out)) throw Error(Status::InvalidArg, "invalid output");
        anneallib::Rng rng(raw_seed);
        const auto s = generateRandomSpins(static_cast<uint64_t>(n), rng);
        std::copy(s.begin(), s.end(), out);
    });
}

rfmf_status rfmf_beta_tables_log(const double* t0s, int64_t chains, int64_t total_steps, int64_t* run_end,
                                 int32_t* run_val, int64_t cap, int64_t* n_runs, int64_t* chain_run_off,
                                 int64_t* n_tables) {
    return guarded([&] {
        if (!t0s || chains <= 0 || !n_runs) throw Error(Status::InvalidArg, "invalid t0 list");
        const BetaRunTables t = buildBetaRunTables(std::vector<double>(t0s, t0s + chains), total_steps);
        *n_runs = static_cast<int64_t>(t.runEnd.size());
        if (n_tables) *n_tables = t.nTables;
        if (chain_run_off) std::copy(t.chainRunOff.begin(), t.chainRunOff.end(), chain_run_off);
        if (run_end || run_val) {
            if (cap < *n_runs) throw Error(Status::InvalidArg, "run buffers smaller than the run count");
            if (run_end) std::copy(t.runEnd.begin(), t.runEnd.end(), run_end);
            if (run_val) std::copy(t.runVal.begin(), t.runVal.end(), run_val);
        }
    });
}

rfmf_status rfmf_beta_tables_linear(int32_t b0, int32_t b1, int64_t total_steps, int64_t* run_end,
                                    int32_t* run_val, int64_t cap, int64_t* n_runs) {
    return guarded([&] {
        if (!n_runs) throw Error(Status::InvalidArg, "null run count");
        const BetaRunTables t = buildLinearBetaRunTable(b0, b1, total_steps, 1);
        *n_runs = static_cast<int64_t>(t.runEnd.size());
        if (run_end || run_val) {
            if (cap < *n_runs) throw Error(Status::InvalidArg, "run buffers smaller than the run count");
            if (run_end) std::copy(t.runEnd.begin(), t.runEnd.end(), run_end);
            if (run_val) std::copy(t.runVal.begin(), t.runVal.end(), run_val);
        }
    });
}

// ---------------------------------------------------------------- one-shot

namespace {

void annealOneShot(const rfmf_csr* g, int64_t B, const int8_t* init, int64_t C,
                   const int64_t* runEnd, const int32_t* runVal, const int64_t* runOff, int64_t nRuns,
                   int32_t gateBetaMax, const uint64_t* phases, const uint64_t* iters, int64_t nPhases,
                   const uint64_t* seeds, const rfmf_launch_cfg* cfg, int8_t* end, int8_t* best,
                   float* ms, int64_t* bestVals, int64_t* finalVals) {
    rfmf::Context ctx(cfg ? cfg->device : 0);
    const int32_t gid = ctx.uploadGraph(g->offsets, g->edge_to, g->edge_weight, g->n, g->m2);
    rfmf::JobSpec spec;
    spec.graph = ctx.graph(gid);
    spec.chains = C;
    spec.initialSpins = init;
    spec.seeds = seeds;
    spec.runEnd = runEnd;
    spec.runVal = runVal;
    spec.chainRunOff = runOff;
    spec.nRuns = nRuns;
    spec.phases.assign(phases, phases + nPhases);
    spec.iters.assign(iters, iters + nPhases);
    spec.gateB = B;
    spec.gateBetaMax = gateBetaMax;
    spec.endSpins = end;
    spec.bestSpins = best;
    spec.bestVals = bestVals;
    spec.finalVals = finalVals;
    auto job = ctx.admit(spec, ctx.stream());
    // The drop-in path picks its geometry with the autotuner (the trial
    // slices are part of the run, so the result is unchanged).
    rfmf::LaunchKnobs knobs;
    knobs.autotune = true;
    const float t = ctx.runToCompletion({job.get()}, knobs);
    if (ms) *ms = t;
    ctx.collect(*job, spec, ctx.stream());
}

}  // namespace

rfmf_status rfmf_anneal(const rfmf_csr* g, int64_t max_weighted_deg, const int8_t* initial_spins,
                        int64_t chains, const double* t0s, const uint64_t* phases, const uint64_t* iters,
                        int64_t n_phases, const uint64_t* solver_seeds, const rfmf_launch_cfg* cfg,
                        int8_t* end_spins, int8_t* best_spins, float* kernel_ms, int64_t* kernel_best_vals,
                        int64_t* kernel_final_vals) {
    return guarded([&] {
        rfmf::checkOneShot(g, initial_spins, chains, phases, iters, n_phases, solver_seeds);
        if (!t0s) throw Error(Status::InvalidArg, "null t0s");
        if (g->n == 0 || n_phases == 0) {
            rfmf::passThrough(g, initial_spins, chains, end_spins, best_spins, kernel_ms, kernel_best_vals,
                        kernel_final_vals);
            return;
        }
        const std::vector<uint64_t> it(iters, iters + n_phases);
        const int64_t total = annealTotalSteps(it);
        const std::vector<double> t0(t0s, t0s + chains);
        // Tables first: an over-cap request fails before any device work.
        // The int32 gate is judged at the tables' largest beta (the one of
        // min t0, src/anneal_kernels.cu:637-639) with the caller's B; when it
        // fails the chains score in exact int64 instead, with the same result.
        const BetaRunTables beta = buildBetaRunTables(t0, total);
        annealOneShot(g, max_weighted_deg, initial_spins, chains, beta.runEnd.data(), beta.runVal.data(),
                      beta.chainRunOff.data(), static_cast<int64_t>(beta.runEnd.size()), -1, phases, iters,
                      n_phases, solver_seeds, cfg, end_spins, best_spins, kernel_ms, kernel_best_vals,
                      kernel_final_vals);
    });
}

rfmf_status rfmf_anneal_tables(const rfmf_csr* g, int64_t max_weighted_deg, const int8_t* initial_spins,
                               int64_t chains, const rfmf_beta_tables* beta, const uint64_t* phases,
                               const uint64_t* iters, int64_t n_phases, const uint64_t* solver_seeds,
                               const rfmf_launch_cfg* cfg, int8_t* end_spins, int8_t* best_spins,
                               float* kernel_ms, int64_t* kernel_best_vals, int64_t* kernel_final_vals) {
    return guarded([&] {
        rfmf::checkOneShot(g, initial_spins, chains, phases, iters, n_phases, solver_seeds);
        if (!beta || !beta->run_end || !beta->run_val || !beta->chain_run_off || beta->n_runs < 1)
            throw Error(Status::InvalidArg, "invalid beta tables");
        if (g->n == 0 || n_phases == 0) {
            rfmf::passThrough(g, initial_spins, chains, end_spins, best_spins, kernel_ms, kernel_best_vals,
                        kernel_final_vals);
            return;
        }
        annealTotalSteps(iters, n_phases);
        annealOneShot(g, max_weighted_deg, initial_spins, chains, beta->run_end, beta->run_val,
                      beta->chain_run_off, beta->n_runs, -1, phases, iters, n_phases, solver_seeds, cfg,
                      end_spins, best_spins, kernel_ms, kernel_best_vals, kernel_final_vals);
    });
}

// ---------------------------------------------------------------- contexts

rfmf_status rfmf_ctx_create(int32_t device, rfmf_ctx** out) {
    return guarded([&] {
        if (!out) throw Error(Status::InvalidArg, "null output");
        auto c = std::make_unique<rfmf_ctx>();
        c->ctx = std::make_unique<rfmf::Context>(device);
        *out = c.release();
    });
}

void rfmf_ctx_destroy(rfmf_ctx* ctx) { delete ctx; }

rfmf_status rfmf_ctx_upload_graph(rfmf_ctx* ctx, const rfmf_csr* g, int32_t* graph_id) {
    return guarded([&] {
        if (!ctx || !graph_id) throw Error(Status::InvalidArg, "null context or output");
        requireCsr(g);
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        *graph_id = ctx->ctx->uploadGraph(g->offsets, g->edge_to, g->edge_weight, g->n, g->m2);
    });
}

rfmf_status rfmf_ctx_release_graph(rfmf_ctx* ctx, int32_t graph_id) {
    return guarded([&] {
        if (!ctx) throw Error(Status::InvalidArg, "null context");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        ctx->ctx->releaseGraph(graph_id);
    });
}

rfmf_status rfmf_ctx_device_info(rfmf_ctx* ctx, int32_t* cu_count, int64_t* lds_per_block, int64_t* hbm_bytes,
                                 char* name, int32_t name_cap) {
    return guarded([&] {
        if (!ctx) throw Error(Status::InvalidArg, "null context");
        const hipDeviceProp_t& p = ctx->ctx->props();
        if (cu_count) *cu_count = p.multiProcessorCount;
        if (lds_per_block) *lds_per_block = static_cast<int64_t>(p.sharedMemPerBlock);
        if (hbm_bytes) *hbm_bytes = static_cast<int64_t>(p.totalGlobalMem);
        if (name && name_cap > 0) {
            std::strncpy(name, p.gcnArchName, static_cast<size_t>(name_cap - 1));
            name[name_cap - 1] = '\0';
        }
    });
}

rfmf_status rfmf_ctx_last_launch(rfmf_ctx* ctx, int32_t* threads, int32_t* grid, int32_t* word_bits,
                                 int32_t* occupancy, float* trial_ms_per_sweep, int32_t* score_mode,
                                 int32_t* state_in_hbm, int32_t* rows_in_lds) {
    return guarded([&] {
        if (!ctx) throw Error(Status::InvalidArg, "null context");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        const rfmf::LaunchReport& r = ctx->ctx->lastLaunch();
        if (threads) *threads = r.threads;
        if (grid) *grid = r.grid;
        if (word_bits) *word_bits = r.bits;
        if (occupancy) *occupancy = r.occupancy;
        if (score_mode) *score_mode = r.mode;
        if (state_in_hbm) *state_in_hbm = r.hbm ? 1 : 0;
        if (rows_in_lds) *rows_in_lds = r.rows ? 1 : 0;
        if (trial_ms_per_sweep)
            for (int i = 0; i < 4; ++i) trial_ms_per_sweep[i] = r.trialMsPerSweep[i];
    });
}

rfmf_status rfmf_ctx_copy_probe(rfmf_ctx* ctx, int64_t bytes, int32_t iters, float* avg_ms) {
    return guarded([&] {
        if (!ctx || bytes < 16 || iters <= 0 || !avg_ms) throw Error(Status::InvalidArg, "invalid probe");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        rfmf::Context& c = *ctx->ctx;
        const size_t b = static_cast<size_t>(bytes) & ~size_t{15};
        rfmf::DevBuf src(b), dst(b);
        rfmf::hipCheck(hipMemsetAsync(src.get(), 1, b, c.stream()), "hipMemsetAsync");
        hipEvent_t e0, e1;
        rfmf::hipCheck(hipEventCreate(&e0), "hipEventCreate");
        rfmf::hipCheck(hipEventCreate(&e1), "hipEventCreate");
        rfmf::hipCheck(rfmf::launch::streamCopy(src.get(), dst.get(), b, c.stream()), "streamCopy");
        rfmf::hipCheck(hipEventRecord(e0, c.stream()), "hipEventRecord");
        for (int i = 0; i < iters; ++i)
            rfmf::hipCheck(rfmf::launch::streamCopy(src.get(), dst.get(), b, c.stream()), "streamCopy");
        rfmf::hipCheck(hipEventRecord(e1, c.stream()), "hipEventRecord");
        rfmf::hipCheck(hipEventSynchronize(e1), "streamCopy");
        float ms = 0.0f;
        rfmf::hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *avg_ms = ms / static_cast<float>(iters);
    });
}

int64_t rfmf_footprint_bytes(int64_t n, int64_t m2, int64_t chains, int64_t beta_runs, int64_t n_phases) {
    return rfmf::footprintBytes(n, m2, chains, beta_runs, n_phases, 32);   // the widest word below B = 2^30
}

namespace {

rfmf::JobSpec specFromJob(const rfmf_job& j, const rfmf::DevGraph* g) {
    if (j.n_phases < 0 || (j.n_phases > 0 && (!j.phases || !j.iters)))
        throw Error(Status::InvalidArg, "invalid phases / iters");
    rfmf::JobSpec s;
    s.graph = g;
    s.chains = j.chains;
    s.initialSpins = j.initial_spins;
    s.seeds = j.solver_seeds;
    s.runEnd = j.beta.run_end;
    s.runVal = j.beta.run_val;
    s.chainRunOff = j.beta.chain_run_off;
    s.nRuns = j.beta.n_runs;
    s.phases.assign(j.phases, j.phases + j.n_phases);
    s.iters.assign(j.iters, j.iters + j.n_phases);
    annealTotalSteps(s.iters);
    s.endSpins = j.end_spins;
    s.bestSpins = j.best_spins;
    s.bestVals = j.best_vals;
    s.finalVals = j.final_vals;
    return s;
}

rfmf::LaunchKnobs knobsFrom(const rfmf_tuning* t) {
    rfmf::LaunchKnobs k;
    if (t) {
        if (t->slice_sweeps < 0 || t->grid_cap < 0)
            throw Error(Status::InvalidArg, "negative slice length or grid cap");
        if (t->threads_per_chain != 0 && !rfmf::dev::validThreads(t->threads_per_chain))
            throw Error(Status::InvalidArg, "threads_per_chain must be 0, 128, 256, 512 or 1024");
        k.sliceSweeps = t->slice_sweeps;
        k.gridCap = t->grid_cap;
        k.hipGraph = t->use_hip_graph != 0;
        k.threads = t->threads_per_chain;
        k.autotune = t->autotune != 0;
        if (t->score_table < -1 || t->score_table > 1)
            throw Error(Status::InvalidArg, "score_table must be -1, 0 or 1");
        k.scoreTable = t->score_table;
        if (t->state_in_hbm < 0 || t->state_in_hbm > 1)
            throw Error(Status::InvalidArg, "state_in_hbm must be 0 or 1");
        k.stateInHbm = t->state_in_hbm;
        if (t->rows_in_lds < -1 || t->rows_in_lds > 1)
            throw Error(Status::InvalidArg, "rows_in_lds must be -1, 0 or 1");
        k.rowsInLds = t->rows_in_lds;
    }
    return k;
}

}  // namespace

rfmf_status rfmf_batch_run(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs, const rfmf_tuning* tuning,
                           float* kernel_ms) {
    return guarded([&] {
        if (!ctx || n_jobs < 0 || (n_jobs > 0 && !jobs)) throw Error(Status::InvalidArg, "invalid batch");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        rfmf::Context& c = *ctx->ctx;
        std::vector<rfmf::JobSpec> specs;
        specs.reserve(static_cast<size_t>(n_jobs));
        for (int64_t i = 0; i < n_jobs; ++i) specs.push_back(specFromJob(jobs[i], c.graph(jobs[i].graph_id)));
        std::vector<std::unique_ptr<rfmf::ResidentJob>> resident;
        std::vector<rfmf::ResidentJob*> ptrs;
        for (const auto& s : specs) {
            resident.push_back(c.admit(s, c.stream()));
            ptrs.push_back(resident.back().get());
        }
        const float ms = c.runToCompletion(ptrs, knobsFrom(tuning));
        if (kernel_ms) *kernel_ms = ms;
        for (size_t i = 0; i < specs.size(); ++i) c.collect(*resident[i], specs[i], c.stream());
    });
}

}  // extern "C"

struct rfmf_run {
    rfmf_ctx* ctx = nullptr;
    std::vector<rfmf::JobSpec> specs;
    std::vector<std::unique_ptr<rfmf::ResidentJob>> jobs;
    std::vector<rfmf::ResidentJob*> ptrs;
    std::unique_ptr<rfmf::Context::GraphedSlices> graph;
    int64_t graphSlice = 0;
    size_t graphActive = 0;

    size_t activeChains() const {
        size_t a = 0;
        for (const auto* j : ptrs)
            for (uint64_t s : j->stepsDone) a += s < j->totalSteps;
        return a;
    }
};

extern "C" {

rfmf_status rfmf_run_create(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs, rfmf_run** out) {
    return guarded([&] {
        if (!ctx || !out || n_jobs <= 0 || !jobs) throw Error(Status::InvalidArg, "invalid run");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        rfmf::Context& c = *ctx->ctx;
        auto r = std::make_unique<rfmf_run>();
        r->ctx = ctx;
        for (int64_t i = 0; i < n_jobs; ++i) r->specs.push_back(specFromJob(jobs[i], c.graph(jobs[i].graph_id)));
        for (const auto& s : r->specs) {
            r->jobs.push_back(c.admit(s, c.stream()));
            r->ptrs.push_back(r->jobs.back().get());
        }
        *out = r.release();
    });
}

rfmf_status rfmf_run_advance(rfmf_run* run, int64_t sweeps, const rfmf_tuning* tuning, float* kernel_ms,
                             int32_t* finished) {
    return guarded([&] {
        if (!run || sweeps <= 0) throw Error(Status::InvalidArg, "invalid run or sweep count");
        std::lock_guard<std::mutex> lk(run->ctx->ctx->mutex());
        rfmf::Context& c = *run->ctx->ctx;
        const rfmf::LaunchKnobs k = knobsFrom(tuning);
        float ms = 0.0f;
        const size_t active = run->activeChains();
        if (active > 0) {
            if (k.hipGraph) {
                if (!run->graph || run->graphSlice != sweeps || run->graphActive != active) {
                    run->graph = c.captureSlices(run->ptrs, static_cast<uint64_t>(sweeps), k);
                    run->graphSlice = sweeps;
                    run->graphActive = active;
                }
                ms = c.replaySlices(*run->graph, 1);
            } else {
                rfmf::LaunchKnobs s = k;
                s.sliceSweeps = sweeps;
                ms = c.runSlice(run->ptrs, s);
            }
        }
        if (kernel_ms) *kernel_ms = ms;
        if (finished) *finished = run->activeChains() == 0 ? 1 : 0;
    });
}

rfmf_status rfmf_run_collect(rfmf_run* run) {
    return guarded([&] {
        if (!run) throw Error(Status::InvalidArg, "null run");
        std::lock_guard<std::mutex> lk(run->ctx->ctx->mutex());
        rfmf::Context& c = *run->ctx->ctx;
        for (size_t i = 0; i < run->specs.size(); ++i) c.collect(*run->jobs[i], run->specs[i], c.stream());
    });
}

void rfmf_run_destroy(rfmf_run* run) {
    if (!run) return;
    {
        std::lock_guard<std::mutex> lk(run->ctx->ctx->mutex());
        run->graph.reset();
        run->jobs.clear();
    }
    delete run;
}

rfmf_status rfmf_batch_run_windows(rfmf_ctx* ctx, const rfmf_job* jobs, int64_t n_jobs,
                                   const int64_t* step_bounds, int64_t n_windows, const rfmf_tuning* tuning,
                                   float* window_ms) {
    return guarded([&] {
        if (!ctx || n_jobs <= 0 || !jobs || n_windows < 0 || (n_windows > 0 && (!step_bounds || !window_ms)))
            throw Error(Status::InvalidArg, "invalid batch or windows");
        for (int64_t i = 0; i < n_windows; ++i)
            if (step_bounds[i] < 0 || (i > 0 && step_bounds[i] < step_bounds[i - 1]))
                throw Error(Status::InvalidArg, "window bounds must be ascending and non-negative");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        rfmf::Context& c = *ctx->ctx;
        std::vector<rfmf::JobSpec> specs;
        for (int64_t i = 0; i < n_jobs; ++i) specs.push_back(specFromJob(jobs[i], c.graph(jobs[i].graph_id)));
        std::vector<std::unique_ptr<rfmf::ResidentJob>> resident;
        std::vector<rfmf::ResidentJob*> ptrs;
        for (const auto& s : specs) {
            resident.push_back(c.admit(s, c.stream()));
            ptrs.push_back(resident.back().get());
        }
        // Every chain of a fresh batch starts at sweep 0 and each slice
        // advances all of them alike, so every window ends at its bound.
        const rfmf::LaunchKnobs k = knobsFrom(tuning);
        for (int64_t i = 0; i < n_windows; ++i)
            window_ms[i] = c.runUntil(ptrs, k, static_cast<uint64_t>(step_bounds[i]));
        c.runToCompletion(ptrs, k);
        for (size_t i = 0; i < specs.size(); ++i) c.collect(*resident[i], specs[i], c.stream());
    });
}

}  // extern "C"
