// engine.hpp -- device runtime of the annealer: resident graphs, chain
// state in HBM, and the persistent slice launches that advance any set of
// chains (one graph or many) by a number of sweeps.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../host/host_common.hpp"
#include "anneal_device.hpp"
#include "anneal_launch.hpp"

namespace rfmf {

void hipCheck(hipError_t e, const char* what);

// RAII device buffer.  alloc() uses hipMalloc and frees with hipFree, which
// waits for every stream of the device; allocAsync() takes the memory from a
// stream-ordered pool and gives it back with hipFreeAsync on the same stream,
// so retiring a buffer does not stall kernels that are still running.  The
// owner must make sure no queued work on another stream still uses it.
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { take(o); }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { reset(); take(o); }
        return *this;
    }
    void alloc(size_t bytes);
    void allocAsync(size_t bytes, hipMemPool_t pool, hipStream_t stream);
    void reset();
    template <class T> T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    size_t bytes() const { return bytes_; }

private:
    void take(DevBuf& o) {
        p_ = o.p_; bytes_ = o.bytes_; async_ = o.async_; stream_ = o.stream_;
        o.p_ = nullptr; o.bytes_ = 0; o.async_ = false; o.stream_ = nullptr;
    }
    void* p_ = nullptr;
    size_t bytes_ = 0;
    bool async_ = false;
    hipStream_t stream_ = nullptr;
};

// Grow-only pinned host buffer.
class PinnedBuf {
public:
    PinnedBuf() = default;
    ~PinnedBuf();
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    void reserve(size_t bytes);
    void* get() const { return p_; }
    size_t capacity() const { return cap_; }

private:
    void* p_ = nullptr;
    size_t cap_ = 0;
};

// Host side of a graph: the validated CSR (kept for the initial cut), its
// max weighted degree, and the interleaved {to, weight} array the kernels
// read (dropped once uploaded).
struct HostGraph {
    int64_t n = 0;
    int64_t m2 = 0;
    int64_t B = 0;
    int64_t maxDeg = 0;   // longest row (unweighted degree)
    std::vector<uint32_t> offsets;
    std::vector<int32_t> to, w;
    std::vector<int32_t> edges;   // int2[m2]
};
// Validates a CSR (offsets monotone from 0 to m2, neighbour ids in range)
// and builds its HostGraph.  Host-only; any thread.
HostGraph prepareGraph(const uint32_t* offsets, const int32_t* to, const int32_t* w, int64_t n, int64_t m2);
// Compact rows of a graph whose weights are all +1 or -1 and whose n is at
// most 32768: one u16 per directed edge (neighbour id | 0x8000 for weight -1),
// padded to a multiple of 8 entries; empty for any other graph.
std::vector<uint16_t> compactRows(const HostGraph& g);

struct DevGraph {
    DevBuf offsets;   // u32[n+1], padded to 16 B
    DevBuf edges;     // int2[m2] {to, weight}
    DevBuf rows16;    // compact rows u16[m2] padded to 16 B (empty when the graph has none)
    int64_t n = 0;
    int64_t m2 = 0;
    int64_t B = 0;    // max weighted degree
    int64_t maxDeg = 0;   // longest row
    HostGraph host;   // host copy for the initial cut
};

// Host description of one job (a set of chains of one graph).
struct JobSpec {
    const DevGraph* graph = nullptr;
    int64_t chains = 0;
    const int8_t* initialSpins = nullptr;   // chains x n
    const uint64_t* seeds = nullptr;         // chains
    const int64_t* runEnd = nullptr;
    const int32_t* runVal = nullptr;
    const int64_t* chainRunOff = nullptr;    // chains
    int64_t nRuns = 0;
    std::vector<uint64_t> phases, iters;     // as given (P == 0 phases included)
    int64_t gateB = -1;                      // B for the int32 gate (-1: graph's B)
    int32_t gateBetaMax = -1;                // |beta| max for the gate (-1: from tables)
    int8_t* endSpins = nullptr;
    int8_t* bestSpins = nullptr;
    int64_t* bestVals = nullptr;
    int64_t* finalVals = nullptr;
};

// Host side of an admission: everything admit() derives from a JobSpec and
// its graph before any device work (validation, gate, word width, executed
// phases, per-chain beta-table extents, initial cut per chain).  The spec's
// arrays are borrowed and must outlive the placement.
struct HostJob {
    int64_t chains = 0;
    int64_t n = 0;
    int bits = 32;
    int mode = dev::kScore32;   // score arithmetic (dev::ScoreMode)
    int maxP = 1;
    uint64_t totalSteps = 0;
    std::vector<uint64_t> phaseEnd;
    std::vector<uint32_t> phaseP;
    std::vector<int64_t> tableStart;   // per chain
    std::vector<int32_t> tableLen;     // per chain
    std::vector<dev::ChainState> state;
    const int8_t* initialSpins = nullptr;
    const uint64_t* seeds = nullptr;
    const int64_t* runEnd = nullptr;
    const int32_t* runVal = nullptr;
    int64_t nRuns = 0;
};

// Chain state of one admitted job, resident in HBM.
struct ResidentJob {
    const DevGraph* graph = nullptr;
    int64_t chains = 0;
    int bits = 32;
    int mode = dev::kScore32;   // score arithmetic (dev::ScoreMode)
    int maxP = 1;
    uint64_t totalSteps = 0;   // executed sweeps per chain
    DevBuf words, best, states, runEnd, runVal, phaseEnd, phaseP;
    int64_t wordStride = 0, bestStride = 0;
    std::vector<uint64_t> hPhaseEnd;     // executed phases (host copy): cumulative end sweep
    std::vector<uint32_t> hPhaseP;       //   and group count
    std::vector<dev::ChainItem> items;   // one per chain (device pointers filled)
    std::vector<uint64_t> stepsDone;     // host mirror of each chain's step after every enqueued slice
    int64_t footprint = 0;
};

struct LaunchKnobs {
    int64_t sliceSweeps = 0;   // 0: run to completion in one launch
    int gridCap = 0;           // 0: occupancy x CUs
    bool hipGraph = false;
    int threads = 0;           // threads per chain (128/256/512/1024); 0: tuned or heuristic
    bool autotune = false;     // time the candidates on the run's first slices
    int scoreTable = 0;        // 8-bit words: 0 table when it costs no resident chains, 1 always, -1 never
    int rowsInLds = 0;         // compact rows in LDS: 0 when they cost no resident chains, 1 whenever they fit, -1 never
    int stateInHbm = 0;        // 0: only chains whose state exceeds LDS keep it in HBM; 1: every chain
    // Set by the planner, not by callers: the register-rich build is not
    // taken (a launch of several groups, and the autotuner's trials for one).
    bool noRich = false;
    // Tiles the autotuner chose per launch group (Context::groupKey); a group
    // not listed uses `threads`, else the cached choice or the heuristic.
    std::vector<std::pair<uint64_t, int>> groupThreads;
};

// Launch shape of one group of chains that share a kernel variant: tile,
// LDS plan, score arithmetic, build and grid.
struct GroupShape {
    int bits = 8;
    int mode = dev::kScore32;
    int threads = dev::kThreads;
    dev::LdsPlan plan{};
    int grid = 1;
    int occupancy = 1;   // resident workgroups per CU
    int maxP = 1;
    bool hbm = false;
    bool rich = false;   // 24-bit-score build for 4 workgroups per CU (launch::anneal)
    bool rows = false;   // plan.rowsOn: graphs with compact rows read them from LDS
};

class Context;

// Whether a workgroup of `threads` threads holds the LDS plan of a chain of
// maxN vertices (state in LDS unless `hbm`) plus the variant's static LDS.
bool fitsLds(const hipDeviceProp_t& props, int64_t maxN, int bits, int mode, int maxP, int threads, bool hbm);
// Threads per chain from chains per CU and n (no tuning): with several
// chains per CU the 4-wave tile (2-wave for small graphs and HBM-resident
// state), else 8 waves (16 with state in HBM), halved until the plan fits.
int heuristicThreads(const hipDeviceProp_t& props, int bits, int mode, int64_t maxN, int64_t chains, int maxP,
                     bool hbm);
// Completes a group whose bits, mode, hbm, maxP and threads are set: LDS
// plan, the score-table, register-rich and rows-in-LDS choices, occupancy and
// grid for nItems chains.  maxM2: the largest 2m of the group's graphs when
// every one has compact rows, else -1.  Occupancies come from `ctx`'s cache
// when it is given, else from the runtime.  Throws Error(InvalidArg /
// Unsupported).
void shapeGroup(Context* ctx, const hipDeviceProp_t& props, GroupShape& g, int64_t maxN, int64_t maxM2, int64_t nItems,
                const LaunchKnobs& knobs);

// What the last launch of a context used (for reports).
struct LaunchReport {
    int threads = 0;
    int grid = 0;
    int bits = 0;
    int occupancy = 0;
    int mode = 0;   // dev::ScoreMode of the last group
    bool hbm = false;   // the last group's state stayed in HBM
    bool rows = false;  // the last group read its graph's rows from LDS
    bool tuned = false;
    float trialMsPerSweep[4] = {0, 0, 0, 0};   // by kThreadChoices, 0 if not tried
};

class Context {
public:
    explicit Context(int device);
    ~Context();

    int device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    hipStream_t sideStream() const { return side_; }
    const hipDeviceProp_t& props() const { return props_; }

    // Copies a CSR to the device.  With `async` the buffers come from the
    // stream-ordered pool on that stream (the call still returns after the
    // copies have landed) and are freed without a device-wide wait.
    int32_t uploadGraph(const uint32_t* offsets, const int32_t* to, const int32_t* w, int64_t n,
                        int64_t m2, hipStream_t async = nullptr);
    int32_t uploadGraph(HostGraph&& g, hipStream_t async = nullptr);
    // Frees a graph.  `wait` first drains stream(); a caller that passes false
    // must know that no queued launch still reads the graph.
    void releaseGraph(int32_t id, bool wait = true);
    const DevGraph* graph(int32_t id) const;

    // Allocates and initialises a job's chain state on `stream` (initial
    // words from the spins, state {cut, cut, 0, 0}).  Validates the job.
    // With `streamOrdered` every buffer comes from the stream-ordered pool on
    // `stream`, so admitting and retiring jobs never waits for running slices.
    std::unique_ptr<ResidentJob> admit(const JobSpec& spec, hipStream_t stream, bool streamOrdered = false);
    // admit() in two halves: the host half (any thread, no device calls) and
    // the device half (allocation, uploads and the initial-word kernel).
    HostJob prepareJob(const JobSpec& spec, const HostGraph& g) const;
    std::unique_ptr<ResidentJob> placeJob(const HostJob& hj, const DevGraph* g, hipStream_t stream,
                                          bool streamOrdered);
    // Advances every unfinished chain of `jobs` by up to `sliceSweeps` sweeps
    // (all of them when 0), as few launches as there are kernel variants.
    // Returns the kernel-only milliseconds measured with events on stream().
    float runSlice(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs);
    // The same, split: enqueue the launches, then wait for them (the host can
    // admit new work on the side stream in between, and may enqueue the next
    // slice before waiting: launches of one context run in order on stream()).
    // The host mirror of every chain's step (ResidentJob::stepsDone) advances
    // when the slice is enqueued.
    struct SliceHandle {
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        bool empty = true;
    };
    SliceHandle launchSlice(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs);
    float waitSlice(SliceHandle& h);
    // Runs jobs to completion (slices repeated) and returns kernel ms.  With
    // autotuning, chains that share one phase list and sweep are tuned per
    // phase.
    float runToCompletion(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs);
    // The same, stopping every chain at sweep `until` (chains that start a
    // window together stay together: each slice advances all of them alike).
    float runUntil(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs, uint64_t until);
    // Every chain of `jobs` has the same phase list and the same sweep count.
    static bool lockstep(const std::vector<ResidentJob*>& jobs);

    // A fixed set of chains whose slice launches are captured once into a
    // hipGraph (item arrays uploaded once) and replayed per slice.
    struct GraphedSlices {
        std::vector<DevBuf> itemBufs;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<ResidentJob*> jobs;
        uint64_t slice = 0;
        ~GraphedSlices();
    };
    std::unique_ptr<GraphedSlices> captureSlices(const std::vector<ResidentJob*>& jobs, uint64_t sliceSweeps,
                                                 const LaunchKnobs& knobs);
    // Autotuner: plans the launch, then for each launch group (kernel variant
    // and size class) times each thread-count candidate on one short slice of
    // that group's own chains (output-invariant), caches the fastest and
    // records it in knobs->groupThreads; the trial launches' kernel ms are
    // added to *spentMs.  `keyP` (0: the group's largest P) and `budget` (0:
    // the shortest remaining run) narrow the tuning to one phase: the cache
    // key uses keyP and each group's trials stay within a twenty-fifth of
    // the budget.
    void autotune(const std::vector<ResidentJob*>& jobs, LaunchKnobs* knobs, float* spentMs, int keyP = 0,
                  uint64_t budget = 0);
    // Replays the captured slice `times` times; returns kernel-side ms.
    float replaySlices(GraphedSlices& g, int times);
    // Copies a finished job's results to the spec's outputs.
    void collect(const ResidentJob& job, const JobSpec& spec, hipStream_t stream);

    // Resident workgroups per CU of a kernel variant (cached).
    int occupancy(int bits, int mode, int threads, size_t lds, bool hbm = false, bool rich = false, bool rows = false);
    // Threads per chain for a launch shape: the autotuner's cached choice,
    // else a heuristic from chains per CU and n.
    int chooseThreads(int bits, int mode, int64_t maxN, int64_t chains, int maxP, bool hbm = false) const;
    // Whether a chain of n vertices keeps its state in HBM: when its LDS plan
    // does not fit even the smallest tile, or when the knobs ask for it.
    bool stateInHbm(int64_t n, int bits, int mode, int maxP, const LaunchKnobs& knobs) const;
    const LaunchReport& lastLaunch() const { return last_; }

    std::mutex& mutex() { return mu_; }

private:
    struct LaunchGroup : GroupShape {
        std::vector<dev::ChainItem> items;
        uint64_t key = 0;    // groupKey: variant and size class
    };
    // Identity of a launch group across plans: word bits, score mode before
    // the table choice, state in HBM, and the size class of its largest n.
    static uint64_t groupKey(int bits, int mode, bool hbm, int64_t topN, bool split);
    // `slice`: sweeps per launch (~0: to completion).
    std::vector<LaunchGroup> planGroups(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs,
                                        uint64_t slice);
    void finishGroup(LaunchGroup& g, int64_t maxN, int maxP, const LaunchKnobs& knobs);
    // Enqueues the groups of one slice (items[i]: group i's device item array)
    // on the launch lanes; returns the first error.
    hipError_t enqueueGroups(const std::vector<LaunchGroup>& groups, const std::vector<dev::ChainItem*>& items,
                             uint64_t slice);
    static uint64_t tuneKey(int bits, int mode, int64_t maxN, int64_t chains, int maxP, bool hbm);
    bool fitsLds(int64_t maxN, int bits, int mode, int maxP, int threads, bool hbm = false) const {
        return rfmf::fitsLds(props_, maxN, bits, mode, maxP, threads, hbm);
    }
    std::unordered_map<uint64_t, int> tuned_;
    LaunchReport last_;
    // RFMF_PROGRESS set (non-empty, not "0"): runUntil reports each finished
    // slice on stderr, so hours-long runs show how far they are.
    bool progress_ = false;
    void reportProgress(const std::vector<ResidentJob*>& jobs, float kernelMs) const;

    int device_;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;
    // Launch lanes: the groups of one slice run concurrently, lane 0 being
    // stream_ and the others forked from it and joined back (fork_/join_).
    // Three lanes: with the side stream that makes the four hardware queues
    // of a process, so no lane shares a queue (and its order) with another.
    static constexpr int kLanes = 3;
    static constexpr size_t kMaxGroups = 128;
    // Size classes only for launches of at least this many sweeps.
    static constexpr uint64_t kClassMinSweeps = 4096;
    hipStream_t lanes_[kLanes - 1] = {};
    hipEvent_t fork_ = nullptr;
    hipEvent_t join_[kLanes - 1] = {};
    hipDeviceProp_t props_{};
    DevBuf sched_;   // kMaxGroups pairs of u32 work counters, zero between launches
    std::unordered_map<int32_t, std::unique_ptr<DevGraph>> graphs_;
    // Item arrays of enqueued slices: two slots used in turn, each a device
    // buffer, its pinned staging copy and the event after the launches that
    // read it (a slot is reused once that event has completed).
    struct ItemSlot {
        DevBuf dev;
        PinnedBuf host;
        hipEvent_t done = nullptr;
    };
    ItemSlot slots_[2];
    int nextSlot_ = 0;
    hipMemPool_t pool_ = nullptr;
    hipMemPool_t pool();
    int32_t nextGraph_ = 1;
    std::unordered_map<uint64_t, int> occCache_;
    std::mutex mu_;
};

int wordBitsFor(int64_t B);   // 8 / 16 / 32 / 64, or 0 if unsupported

}  // namespace rfmf
