#include "host_common.hpp"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <limits>

using rfmf::Error;
using rfmf::Status;

int64_t annealTotalSteps(const std::vector<uint64_t>& iters) {
    return annealTotalSteps(iters.data(), static_cast<int64_t>(iters.size()));
}

int64_t annealTotalSteps(const uint64_t* iters, int64_t count) {
    uint64_t total = 0;
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t it = iters[i];
        if (it > static_cast<uint64_t>(kMaxTotalSteps) ||
            total + it > static_cast<uint64_t>(kMaxTotalSteps))
            throw Error(Status::StepCap, "Error: --iters totals more than " +
                                             std::to_string(kMaxTotalSteps) +
                                             " sweeps (the total-sweep sanity cap).");
        total += it;
    }
    return static_cast<int64_t>(total);
}

namespace {

void checkRunCap(size_t runs, const std::string& what) {
    if (static_cast<int64_t>(runs) >= kMaxBetaRuns)
        throw Error(Status::BetaCap, "Error: run-length beta tables exceed " +
                                         std::to_string(kMaxBetaRuns) + " runs (" + what + ").");
}

// Log schedule runs.  beta(s) = round(x(s)), x(s) = ln(1+s)/t0 * 2^20, is
// non-decreasing, and for a run of value v the next run starts at the first
// s with x(s) >= v + 1/2, i.e. near exp((v + 1/2) * t0 / 2^20) - 1.  That
// estimate is then corrected against the exact integer function, so the
// table equals the per-step evaluation.
void appendLogRuns(double t0, int64_t total, std::vector<int64_t>& ends, std::vector<int32_t>& vals) {
    if (total <= 0) {
        ends.push_back(0);
        vals.push_back(0);
        return;
    }
    int64_t s = 1;
    for (;;) {
        checkRunCap(ends.size(), "t0=" + std::to_string(t0));
        const int32_t v = betaFixedForStep(static_cast<uint64_t>(s), t0);
        const double x = std::exp((static_cast<double>(v) + 0.5) * t0 /
                                  static_cast<double>(1 << kGumbelFracBits)) - 1.0;
        int64_t next;
        if (!(x < static_cast<double>(total)))
            next = total + 1;
        else
            next = std::max<int64_t>(s + 1, static_cast<int64_t>(std::ceil(x)));
        while (next > s + 1 && betaFixedForStep(static_cast<uint64_t>(next - 1), t0) > v) --next;
        while (next <= total && betaFixedForStep(static_cast<uint64_t>(next), t0) <= v) ++next;
        if (next > total) {
            ends.push_back(total);
            vals.push_back(v);
            return;
        }
        ends.push_back(next - 1);   // 0-based exclusive end of this run
        vals.push_back(v);
        s = next;
    }
}

}  // namespace

BetaRunTables buildBetaRunTables(const std::vector<double>& t0s, int64_t totalSteps) {
    BetaRunTables t;
    t.chainRunOff.resize(t0s.size());
    std::vector<double> distinct;
    std::vector<int64_t> first;
    for (size_t c = 0; c < t0s.size(); ++c) {
        if (!(t0s[c] > 0.0))
            throw Error(Status::InvalidArg, "Error: every t0 must be > 0");
        const auto it = std::find(distinct.begin(), distinct.end(), t0s[c]);
        if (it == distinct.end()) {
            distinct.push_back(t0s[c]);
            first.push_back(static_cast<int64_t>(t.runEnd.size()));
            appendLogRuns(t0s[c], totalSteps, t.runEnd, t.runVal);
            t.chainRunOff[c] = first.back();
        } else {
            t.chainRunOff[c] = first[static_cast<size_t>(it - distinct.begin())];
        }
    }
    t.nTables = static_cast<int64_t>(distinct.size());
    return t;
}

int32_t linearBetaForStep(uint64_t step, int64_t totalSteps, int32_t b0, int32_t b1) {
    if (totalSteps <= 1) return b0;
    const __int128 num = static_cast<__int128>(static_cast<int64_t>(b1) - b0) *
                         static_cast<__int128>(static_cast<int64_t>(step) - 1);
    return static_cast<int32_t>(b0 + static_cast<int64_t>(num / (totalSteps - 1)));
}

BetaRunTables buildLinearBetaRunTable(int32_t b0, int32_t b1, int64_t totalSteps, int64_t chains) {
    BetaRunTables t;
    t.chainRunOff.assign(static_cast<size_t>(chains), 0);
    t.nTables = 1;
    if (totalSteps <= 0) {
        t.runEnd.push_back(0);
        t.runVal.push_back(0);
        return t;
    }
    // The value is monotone in s, so each run is found by galloping from its
    // first step and bisecting the last step holding the same value.
    auto f = [&](int64_t s) { return linearBetaForStep(static_cast<uint64_t>(s), totalSteps, b0, b1); };
    int64_t s = 1;
    while (s <= totalSteps) {
        checkRunCap(t.runEnd.size(), "linear schedule");
        const int32_t v = f(s);
        int64_t lo = s, span = 1, hi;
        for (;;) {
            hi = std::min(totalSteps + 1, lo + span);
            if (hi > totalSteps || f(hi) != v) break;
            lo = hi;
            span *= 2;
        }
        while (hi - lo > 1) {   // f(lo) == v, and hi is past the run or the table
            const int64_t mid = lo + (hi - lo) / 2;
            if (f(mid) == v) lo = mid; else hi = mid;
        }
        t.runEnd.push_back(lo);   // last 1-based step of the run == 0-based exclusive end
        t.runVal.push_back(v);
        s = lo + 1;
    }
    return t;
}

bool scoreSafeInt32ForBeta(int64_t betaMaxFixed, int64_t deltaBound) {
    const int64_t limit = static_cast<int64_t>(std::numeric_limits<int32_t>::max()) - kGumbelMaxMag;
    if (betaMaxFixed < 0) betaMaxFixed = -betaMaxFixed;
    if (betaMaxFixed > 0 && deltaBound > 0 && deltaBound > limit / betaMaxFixed) return false;
    return betaMaxFixed * deltaBound + kGumbelMaxMag <
           static_cast<int64_t>(std::numeric_limits<int32_t>::max());
}

bool annealScoreSafeInt32(int64_t totalSteps, double t0, int64_t deltaBound) {
    const double betaMax = std::log(1.0 + static_cast<double>(totalSteps)) / t0;
    const int64_t bmf = static_cast<int64_t>(std::llround(betaMax * (1 << kGumbelFracBits)));
    return scoreSafeInt32ForBeta(bmf, deltaBound);
}

int64_t maxWeightedDegree(const anneallib::AnnealCsr& ig) {
    int64_t B = 0;
    const auto& off = ig.offsets();
    const auto& w = ig.edgeWeight();
    for (size_t v = 0; v + 1 < off.size(); ++v) {
        int64_t s = 0;
        for (uint32_t i = off[v]; i < off[v + 1]; ++i) s += std::llabs(static_cast<long long>(w[i]));
        B = std::max(B, s);
    }
    return B;
}

void removeDegreeZeroVertices(anneallib::Graph& g) {
    std::vector<int64_t> isolated;
    for (int64_t v : g.getVertices())
        if (g.degree(v) == 0) isolated.push_back(v);
    for (int64_t v : isolated) g.removeVertex(v);
}

std::vector<int8_t> generateRandomSpins(uint64_t n, anneallib::Rng& rng) {
    std::vector<int8_t> spin(n);
    for (uint64_t i = 0; i < n; ++i) spin[i] = rng.bits(1) ? int8_t{1} : int8_t{-1};
    return spin;
}
