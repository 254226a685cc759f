// host_common.hpp -- host-side inputs of the annealing path: beta schedules,
// the int32 score gate, the weighted-degree bound, degree-zero removal and
// the initial spins (reference: src/host_common.h:22-131; generateRandomSpins,
// removeDegreeZeroVertices, annealTotalSteps, buildBetaRunTables,
// annealScoreSafeInt32 and maxWeightedDegree at
// src/host_common.cpp:40-45,184-325).  Errors are reported as rfmf::Error
// (a status code and a message) instead of terminating the process.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "graph.hpp"
#include "immutable_graph.hpp"
#include "rng.hpp"

namespace rfmf {

// Status codes, identical to the C ABI's rfmf_status values.
enum class Status : int {
    Ok = 0,
    InvalidArg = 1,
    Unsupported = 2,
    Hip = 3,
    NoDevice = 4,
    OutOfMemory = 5,
    BetaCap = 6,
    StepCap = 7,
    Io = 8,
    NotFound = 9,
    Admission = 10,
    State = 11,
};

class Error : public std::runtime_error {
public:
    Error(Status s, const std::string& msg) : std::runtime_error(msg), status_(s) {}
    Status status() const { return status_; }

private:
    Status status_;
};

}  // namespace rfmf

// Solver-stream seed = raw seed + this offset; the initial spins use the raw
// seed (src/host_common.h:22-25).
constexpr uint64_t kSolverSeedOffset = 0xC2B2AE3D27D4EB4FULL;

// Q20 fixed point of every Gumbel, beta and score.
constexpr int kGumbelFracBits = 20;
// |Gumbel| bound: LUT[15] + 127 (the largest coarse quantile plus the 7-bit
// fine offset).
constexpr int64_t kGumbelMaxMag = 3617613;

// beta_fixed of a 1-based sweep: round(ln(1+step)/t0 * 2^20), always on the
// host in double so that every consumer sees the same integer
// (src/host_common.h:70-76).
inline int32_t betaFixedForStep(uint64_t step, double t0) {
    const double beta = std::log(1.0 + static_cast<double>(step)) / t0;
    return static_cast<int32_t>(std::round(beta * (1 << kGumbelFracBits)));
}

// Sanity cap on sweeps per chain (src/host_common.h:78-80).
constexpr int64_t kMaxTotalSteps = 1'000'000'000'000;
// Cap on beta runs over all tables (src/host_common.h:96-99).
constexpr int64_t kMaxBetaRuns = 33'000'000;

// Run-length beta tables (src/host_common.h:82-94): run r of a table covers
// 0-based steps [runEnd[r-1], runEnd[r]) with value runVal[r]; the last end is
// the schedule length; every table has at least one run.
struct BetaRunTables {
    std::vector<int64_t> runEnd;
    std::vector<int32_t> runVal;
    std::vector<int64_t> chainRunOff;   // chain -> first run of its table
    int64_t nTables = 0;
};

// Sum of per-phase sweep counts, checked against kMaxTotalSteps in unsigned
// arithmetic per entry.  Throws rfmf::Error(StepCap).
int64_t annealTotalSteps(const std::vector<uint64_t>& iters);
// The same over `count` entries (no copies).
int64_t annealTotalSteps(const uint64_t* iters, int64_t count);

// One log table per distinct t0 (exact double equality), over 1-based sweeps
// [1, totalSteps].  Throws rfmf::Error(BetaCap / InvalidArg).
BetaRunTables buildBetaRunTables(const std::vector<double>& t0s, int64_t totalSteps);

// Exact-integer linear schedule over 1-based sweeps [1, totalSteps]:
// beta(s) = b0 + ((b1 - b0) * (s - 1)) / (totalSteps - 1), Q20, shared by every
// chain (chainRunOff all zero).  Throws rfmf::Error(BetaCap).
BetaRunTables buildLinearBetaRunTable(int32_t b0, int32_t b1, int64_t totalSteps, int64_t chains);
int32_t linearBetaForStep(uint64_t step, int64_t totalSteps, int32_t b0, int32_t b1);

// True iff the packed int32 score path is safe: round(ln(1+T)/t0 * 2^20) * B +
// |Gumbel|max < INT32_MAX, with the product guarded against int64 overflow
// (src/host_common.cpp:296-312).
bool annealScoreSafeInt32(int64_t totalSteps, double t0, int64_t deltaBound);
// Same verdict from an explicit largest |beta| (used for arbitrary tables).
bool scoreSafeInt32ForBeta(int64_t betaMaxFixed, int64_t deltaBound);

// B = max_v sum |w| over v's row (src/host_common.cpp:314-325).
int64_t maxWeightedDegree(const anneallib::AnnealCsr& ig);

// Removes every degree-zero vertex in place; must run before AnnealCsr(g)
// because it changes the dense id space (src/host_common.cpp:184-194).
void removeDegreeZeroVertices(anneallib::Graph& g);

// spin[i] = +1 if the next reservoir bit is set, else -1, i = 0..n-1, from the
// generator seeded with the RAW seed (src/host_common.cpp:40-45).
std::vector<int8_t> generateRandomSpins(uint64_t n, anneallib::Rng& rng);
