// tune_flips.cpp -- quality-gated choice of the flips per sweep (the
// group-count ladder P, P/2, ..., 1).
//
// No reference counterpart: the reference takes the ladder as an input
// (main.cu --phases) and its committed runs all use 128..1.  More groups mean
// more flips per sweep; whether that buys a better cut for a given sweep
// budget depends on the graph, so the tuner measures it.  Each candidate
// ladder anneals the same pilot chains (same raw seeds, hence the same initial
// spins and Gumbel streams) over the full budget; a candidate is eligible when
// its median best cut is within a relative tolerance of a floor median (a
// reference ladder's, e.g. the reference's 128..1, or else the best one's),
// and the eligible candidate with the highest throughput wins.  Pilot runs go
// through the ordinary batch path, so their geometry is autotuned as well.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../../include/rfmf.h"
#include "../host/host_common.hpp"
#include "../host/rng.hpp"
#include "capi_internal.hpp"

using rfmf::Error;
using rfmf::Status;

namespace {

std::vector<uint64_t> ladderOf(uint64_t top) {
    if (top == 0 || (top & (top - 1)) != 0)
        throw Error(Status::InvalidArg, "a top group count must be a power of two");
    if (top > static_cast<uint64_t>(rfmf::dev::kMaxGroups))
        throw Error(Status::Unsupported, "top group count above the 4096-group limit");
    std::vector<uint64_t> p;
    for (uint64_t x = top;; x /= 2) {
        p.push_back(x);
        if (x == 1) break;
    }
    return p;
}

std::vector<uint64_t> itersOf(size_t len, int64_t sweeps) {
    std::vector<uint64_t> it(len, static_cast<uint64_t>(sweeps) / len);
    for (size_t k = 0; k < static_cast<uint64_t>(sweeps) % len; ++k) ++it[k];
    return it;
}

double median(std::vector<int64_t> v) {
    std::sort(v.begin(), v.end());
    const size_t m = v.size() / 2;
    return v.size() % 2 ? static_cast<double>(v[m]) : 0.5 * (static_cast<double>(v[m - 1]) + v[m]);
}

}  // namespace

extern "C" {

rfmf_status rfmf_flips_ladder(uint64_t top_p, int64_t sweeps, uint64_t* phases, uint64_t* iters, int64_t cap,
                              int64_t* n_phases) {
    return rfmf::guarded([&] {
        if (sweeps < 0 || !n_phases) throw Error(Status::InvalidArg, "negative sweeps or null output");
        const std::vector<uint64_t> p = ladderOf(top_p);
        *n_phases = static_cast<int64_t>(p.size());
        if (!phases && !iters) return;
        if (!phases || !iters || cap < static_cast<int64_t>(p.size()))
            throw Error(Status::InvalidArg, "ladder output arrays missing or too small");
        const std::vector<uint64_t> it = itersOf(p.size(), sweeps);
        std::copy(p.begin(), p.end(), phases);
        std::copy(it.begin(), it.end(), iters);
    });
}

rfmf_status rfmf_tune_flips(rfmf_ctx* ctx, int32_t graph_id, const rfmf_flips_tune* spec,
                            rfmf_flips_report* report) {
    return rfmf::guarded([&] {
        if (!ctx || !spec || !report || !spec->top_p || spec->n_candidates < 1 ||
            spec->n_candidates > RFMF_MAX_FLIPS_CANDIDATES || spec->sweeps < 1 || spec->pilot_chains < 0 ||
            spec->reference < -1 || spec->reference >= spec->n_candidates || !(spec->quality_tol >= 0.0))
            throw Error(Status::InvalidArg, "invalid flips-tuning request");
        std::lock_guard<std::mutex> lk(ctx->ctx->mutex());
        rfmf::Context& c = *ctx->ctx;
        const rfmf::DevGraph* g = c.graph(graph_id);
        const int64_t C = spec->pilot_chains > 0 ? spec->pilot_chains : 256;
        const int64_t n = g->n;

        // Pilot inputs shared by every candidate: the production rules for
        // initial spins (raw seed) and solver seeds, one linear beta table.
        std::vector<int8_t> init(static_cast<size_t>(C * n));
        std::vector<uint64_t> seeds(static_cast<size_t>(C));
        for (int64_t k = 0; k < C; ++k) {
            const uint64_t raw = spec->seed_base + 1 + static_cast<uint64_t>(k);
            anneallib::Rng rng(raw);
            const std::vector<int8_t> s = generateRandomSpins(static_cast<uint64_t>(n), rng);
            std::copy(s.begin(), s.end(), init.begin() + k * n);
            seeds[k] = raw + kSolverSeedOffset;
        }
        const BetaRunTables beta =
            buildLinearBetaRunTable(spec->beta0_q20, spec->beta1_q20, spec->sweeps, C);
        std::vector<int64_t> bestVals(static_cast<size_t>(C));

        *report = rfmf_flips_report{};
        report->n_candidates = spec->n_candidates;
        for (int32_t k = 0; k < spec->n_candidates; ++k) {
            const std::vector<uint64_t> phases = ladderOf(spec->top_p[k]);
            rfmf::JobSpec js;
            js.graph = g;
            js.chains = C;
            js.initialSpins = init.data();
            js.seeds = seeds.data();
            js.runEnd = beta.runEnd.data();
            js.runVal = beta.runVal.data();
            js.chainRunOff = beta.chainRunOff.data();
            js.nRuns = static_cast<int64_t>(beta.runEnd.size());
            js.phases = phases;
            js.iters = itersOf(phases.size(), spec->sweeps);
            js.bestVals = bestVals.data();
            auto job = c.admit(js, c.stream());
            rfmf::LaunchKnobs knobs;
            knobs.autotune = true;
            const float ms = c.runToCompletion({job.get()}, knobs);
            c.collect(*job, js, c.stream());
            report->median_best[k] = median(bestVals);
            report->chain_steps_per_s[k] =
                ms > 0.0f ? static_cast<double>(C) * static_cast<double>(spec->sweeps) / (ms * 1e-3) : 0.0;
        }
        double top = report->median_best[0];
        for (int32_t k = 1; k < spec->n_candidates; ++k) top = std::max(top, report->median_best[k]);
        const double anchor = spec->reference >= 0 ? report->median_best[spec->reference] : top;
        const double floorQ = anchor - spec->quality_tol * std::fabs(anchor);
        int32_t chosen = -1;
        for (int32_t k = 0; k < spec->n_candidates; ++k) {
            if (report->median_best[k] < floorQ) continue;
            if (chosen < 0 || report->chain_steps_per_s[k] > report->chain_steps_per_s[chosen]) chosen = k;
        }
        report->chosen = chosen;
    });
}

}  // extern "C"
