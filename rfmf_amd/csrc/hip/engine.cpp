// engine.cpp -- device runtime: graph upload, chain-state admission, the
// slice launches and result collection.
//
// Host-side duties follow the reference driver (src/anneal_kernels.cu:611-831):
// CSR flattening, the int32 gate verdict, initial cut per chain with the host
// cut recompute, kernel-only event timing.  Unlike the reference, device
// buffers are owned by a context and kept resident, chains of many graphs go
// into one persistent launch, and a run can be cut into slices.
#include "engine.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_set>

#include "../host/max_cut.hpp"
#include "../host/tuning.hpp"

namespace rfmf {

void hipCheck(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(Status::OutOfMemory, std::string(what) + ": " + hipGetErrorString(e));
    throw Error(Status::Hip, std::string(what) + ": " + hipGetErrorString(e));
}

void DevBuf::alloc(size_t bytes) {
    reset();
    if (bytes == 0) bytes = 16;
    hipCheck(hipMalloc(&p_, bytes), "hipMalloc");
    bytes_ = bytes;
}

void DevBuf::allocAsync(size_t bytes, hipMemPool_t pool, hipStream_t stream) {
    reset();
    if (bytes == 0) bytes = 16;
    hipCheck(hipMallocFromPoolAsync(&p_, bytes, pool, stream), "hipMallocFromPoolAsync");
    bytes_ = bytes;
    async_ = true;
    stream_ = stream;
}

void DevBuf::reset() {
    if (p_) (void)(async_ ? hipFreeAsync(p_, stream_) : hipFree(p_));
    p_ = nullptr;
    bytes_ = 0;
    async_ = false;
    stream_ = nullptr;
}

PinnedBuf::~PinnedBuf() {
    if (p_) (void)hipHostFree(p_);
}

void PinnedBuf::reserve(size_t bytes) {
    if (bytes <= cap_) return;
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    cap_ = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 2, 64 << 10);
    hipCheck(hipHostMalloc(&p_, want, hipHostMallocDefault), "hipHostMalloc");
    cap_ = want;
}

int wordBitsFor(int64_t B) {
    if (B <= dev::WordFmt<8>::kMaxB) return 8;
    if (B <= dev::WordFmt<16>::kMaxB) return 16;
    if (B <= dev::WordFmt<32>::kMaxB) return 32;
    if (B <= dev::WordFmt<64>::kMaxB) return 64;
    return 0;
}

Context::Context(int device) : device_(device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw Error(Status::NoDevice, "no HIP device visible");
    if (device < 0 || device >= count)
        throw Error(Status::NoDevice, "device ordinal " + std::to_string(device) + " out of range");
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    hipCheck(hipGetDeviceProperties(&props_, device_), "hipGetDeviceProperties");
    hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hipCheck(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "hipStreamCreate");
    for (hipStream_t& l : lanes_) hipCheck(hipStreamCreateWithFlags(&l, hipStreamNonBlocking), "hipStreamCreate");
    hipCheck(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
    for (hipEvent_t& e : join_) hipCheck(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    sched_.alloc(2 * kMaxGroups * sizeof(uint32_t));
    hipCheck(hipMemset(sched_.get(), 0, sched_.bytes()), "hipMemset");
    const char* pr = std::getenv("RFMF_PROGRESS");
    progress_ = pr && *pr && std::strcmp(pr, "0") != 0;
}

void Context::reportProgress(const std::vector<ResidentJob*>& jobs, float kernelMs) const {
    uint64_t done = ~uint64_t{0}, total = 0;
    size_t chains = 0;
    for (const ResidentJob* j : jobs) {
        total = std::max(total, j->totalSteps);
        chains += j->stepsDone.size();
        for (uint64_t s : j->stepsDone) done = std::min(done, s);
    }
    if (chains == 0) return;
    std::fprintf(stderr, "rfmf: %zu chains at sweep %llu of %llu, %.1f s of kernel time\n", chains,
                 static_cast<unsigned long long>(done), static_cast<unsigned long long>(total), kernelMs / 1e3);
    std::fflush(stderr);
}

Context::~Context() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    for (hipStream_t l : lanes_)
        if (l) (void)hipStreamSynchronize(l);
    graphs_.clear();
    sched_.reset();
    for (ItemSlot& s : slots_) {
        if (s.done) (void)hipEventDestroy(s.done);
        s.dev.reset();
    }
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_) (void)hipStreamSynchronize(side_);
    if (pool_) (void)hipMemPoolDestroy(pool_);
    for (hipStream_t l : lanes_)
        if (l) (void)hipStreamDestroy(l);
    if (fork_) (void)hipEventDestroy(fork_);
    for (hipEvent_t e : join_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (side_) (void)hipStreamDestroy(side_);
}

hipMemPool_t Context::pool() {
    if (pool_) return pool_;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = device_;
    hipCheck(hipMemPoolCreate(&pool_, &props), "hipMemPoolCreate");
    // Keep freed memory in the pool: a serving loop admits and retires jobs
    // all the time, and returning memory to the driver costs a device sync.
    uint64_t keep = ~uint64_t{0};
    hipCheck(hipMemPoolSetAttribute(pool_, hipMemPoolAttrReleaseThreshold, &keep), "hipMemPoolSetAttribute");
    return pool_;
}

HostGraph prepareGraph(const uint32_t* offsets, const int32_t* to, const int32_t* w, int64_t n, int64_t m2) {
    if (!offsets || n < 0 || m2 < 0 || (m2 > 0 && (!to || !w)))
        throw Error(Status::InvalidArg, "uploadGraph: null CSR array or negative size");
    if (n >= (int64_t{1} << 31) || m2 >= (int64_t{1} << 32))
        throw Error(Status::Unsupported, "uploadGraph: graph exceeds 32-bit ids/offsets");
    if (offsets[0] != 0 || static_cast<int64_t>(offsets[n]) != m2)
        throw Error(Status::InvalidArg, "uploadGraph: offsets[0] must be 0 and offsets[n] == m2");
    HostGraph g;
    g.n = n;
    g.m2 = m2;
    int64_t B = 0, maxDeg = 0;
    for (int64_t v = 0; v < n; ++v) {
        if (offsets[v + 1] < offsets[v]) throw Error(Status::InvalidArg, "uploadGraph: offsets not monotone");
        maxDeg = std::max<int64_t>(maxDeg, offsets[v + 1] - offsets[v]);
        int64_t s = 0;
        for (uint32_t i = offsets[v]; i < offsets[v + 1]; ++i) {
            if (to[i] < 0 || to[i] >= n) throw Error(Status::InvalidArg, "uploadGraph: neighbour id out of range");
            s += std::llabs(static_cast<long long>(w[i]));
        }
        B = std::max(B, s);
    }
    g.B = B;
    g.maxDeg = maxDeg;
    g.offsets.assign(offsets, offsets + n + 1);
    g.to.assign(to, to + m2);
    g.w.assign(w, w + m2);
    g.edges.resize(static_cast<size_t>(2 * m2));
    for (int64_t i = 0; i < m2; ++i) {
        g.edges[2 * i] = to[i];
        g.edges[2 * i + 1] = w[i];
    }
    return g;
}

int32_t Context::uploadGraph(const uint32_t* offsets, const int32_t* to, const int32_t* w,
                             int64_t n, int64_t m2, hipStream_t async) {
    return uploadGraph(prepareGraph(offsets, to, w, n, m2), async);
}

std::vector<uint16_t> compactRows(const HostGraph& g) {
    std::vector<uint16_t> rows;
    if (g.m2 <= 0 || g.n > dev::kCompactMaxN) return rows;
    for (int32_t w : g.w)
        if (w != 1 && w != -1) return rows;
    rows.assign(static_cast<size_t>((g.m2 + 7) / 8 * 8), 0);
    for (int64_t e = 0; e < g.m2; ++e)
        rows[e] = static_cast<uint16_t>(static_cast<uint32_t>(g.to[e]) | (g.w[e] < 0 ? dev::kCompactNeg : 0u));
    return rows;
}

int32_t Context::uploadGraph(HostGraph&& hg, hipStream_t async) {
    auto g = std::make_unique<DevGraph>();
    g->n = hg.n;
    g->m2 = hg.m2;
    g->B = hg.B;
    g->maxDeg = hg.maxDeg;
    g->host = std::move(hg);
    const int64_t n = g->n, m2 = g->m2;
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    // Offsets padded to 16 B (kernels copy them into LDS in 16-byte pieces).
    const size_t offBytes = static_cast<size_t>(n + 1) * sizeof(uint32_t);
    const size_t offAlloc = (offBytes + 15) / 16 * 16;
    const size_t edgeBytes = static_cast<size_t>(std::max<int64_t>(m2, 1)) * 8;
    const std::vector<uint16_t> rows = compactRows(g->host);
    const size_t rowBytes = rows.size() * sizeof(uint16_t);
    if (async) {
        g->offsets.allocAsync(offAlloc, pool(), async);
        g->edges.allocAsync(edgeBytes, pool(), async);
        if (rowBytes) g->rows16.allocAsync(rowBytes, pool(), async);
        hipCheck(hipMemcpyAsync(g->offsets.get(), g->host.offsets.data(), offBytes, hipMemcpyHostToDevice, async),
                 "upload offsets");
        if (m2 > 0)
            hipCheck(hipMemcpyAsync(g->edges.get(), g->host.edges.data(), m2 * 8, hipMemcpyHostToDevice, async),
                     "upload edges");
        if (rowBytes)
            hipCheck(hipMemcpyAsync(g->rows16.get(), rows.data(), rowBytes, hipMemcpyHostToDevice, async),
                     "upload compact rows");
        hipCheck(hipStreamSynchronize(async), "upload graph");
    } else {
        g->offsets.alloc(offAlloc);
        g->edges.alloc(edgeBytes);
        if (rowBytes) g->rows16.alloc(rowBytes);
        hipCheck(hipMemcpy(g->offsets.get(), g->host.offsets.data(), offBytes, hipMemcpyHostToDevice),
                 "upload offsets");
        if (m2 > 0)
            hipCheck(hipMemcpy(g->edges.get(), g->host.edges.data(), m2 * 8, hipMemcpyHostToDevice), "upload edges");
        if (rowBytes)
            hipCheck(hipMemcpy(g->rows16.get(), rows.data(), rowBytes, hipMemcpyHostToDevice), "upload compact rows");
    }
    std::vector<int32_t>().swap(g->host.edges);
    const int32_t id = nextGraph_++;
    graphs_.emplace(id, std::move(g));
    return id;
}

void Context::releaseGraph(int32_t id, bool wait) {
    const auto it = graphs_.find(id);
    if (it == graphs_.end()) throw Error(Status::NotFound, "unknown graph id " + std::to_string(id));
    if (wait) hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    graphs_.erase(it);
}

const DevGraph* Context::graph(int32_t id) const {
    const auto it = graphs_.find(id);
    if (it == graphs_.end()) throw Error(Status::NotFound, "unknown graph id " + std::to_string(id));
    return it->second.get();
}

int Context::occupancy(int bits, int mode, int threads, size_t lds, bool hbm, bool rich, bool rows) {
    const uint64_t key = (static_cast<uint64_t>(rows) << 58) | (static_cast<uint64_t>(rich) << 57) |
                         (static_cast<uint64_t>(hbm) << 56) |
                         (static_cast<uint64_t>(bits) << 48) |
                         (static_cast<uint64_t>(mode) << 44) | (static_cast<uint64_t>(threads) << 32) | lds;
    const auto it = occCache_.find(key);
    if (it != occCache_.end()) return it->second;
    const int occ = launch::maxActiveBlocks(bits, mode, threads, hbm, rich, rows, lds);
    occCache_[key] = occ;
    return occ;
}

uint64_t Context::tuneKey(int bits, int mode, int64_t maxN, int64_t chains, int maxP, bool hbm) {
    auto lg = [](int64_t x) { int l = 0; while ((int64_t{1} << (l + 1)) <= x) ++l; return static_cast<uint64_t>(l); };
    return (static_cast<uint64_t>(hbm) << 48) | (static_cast<uint64_t>(bits) << 42) |
           (static_cast<uint64_t>(mode) << 40) |
           (lg(std::max<int64_t>(maxN, 1)) << 32) | (lg(std::max<int64_t>(chains, 1)) << 24) |
           static_cast<uint64_t>(maxP);
}

bool fitsLds(const hipDeviceProp_t& props, int64_t maxN, int bits, int mode, int maxP, int threads, bool hbm) {
    const dev::LdsPlan p = dev::makeLdsPlan(maxN, bits, maxP, threads, mode, hbm);
    return static_cast<size_t>(p.total) + dev::staticLds(mode) <= props.sharedMemPerBlock;
}

bool Context::stateInHbm(int64_t n, int bits, int mode, int maxP, const LaunchKnobs& knobs) const {
    return knobs.stateInHbm > 0 || !fitsLds(n, bits, mode, maxP, 128);
}

int heuristicThreads(const hipDeviceProp_t& props, int bits, int mode, int64_t maxN, int64_t chains, int maxP,
                     bool hbm) {
    // With several chains per CU the 4-wave tile keeps every SIMD busy across
    // chains (the 2-wave tile for small graphs, and for state in HBM, where
    // LDS no longer limits the chains per CU); with about one chain per CU or
    // fewer, 8 waves per chain cut the per-sweep latency (16 waves add more
    // per-wave work than they hide), and 16 when the state is in HBM, whose
    // loads want more waves in flight.
    namespace T = gpu3::tuning;
    const double perCu = static_cast<double>(chains) / std::max(1, props.multiProcessorCount);
    int t = hbm ? T::tpbStateInHbm : maxN <= T::tpbWalkSmallMaxN ? T::tpbWalkSmall : T::tpbWalkLarge;
    if (perCu >= T::coResidentMinChainsPerCu)
        t = (hbm || maxN <= T::blocked8CoResBigMinN) ? T::tpbCoResidentBlocked8 : T::tpbCoResidentBlocked8BigN;
    while (t > 128 && !fitsLds(props, maxN, bits, mode, maxP, t, hbm)) t /= 2;
    return t;
}

int Context::chooseThreads(int bits, int mode, int64_t maxN, int64_t chains, int maxP, bool hbm) const {
    const auto it = tuned_.find(tuneKey(bits, mode, maxN, chains, maxP, hbm));
    if (it != tuned_.end()) return it->second;
    return heuristicThreads(props_, bits, mode, maxN, chains, maxP, hbm);
}

std::unique_ptr<ResidentJob> Context::admit(const JobSpec& spec, hipStream_t stream, bool streamOrdered) {
    if (!spec.graph) throw Error(Status::InvalidArg, "job has no graph");
    return placeJob(prepareJob(spec, spec.graph->host), spec.graph, stream, streamOrdered);
}

HostJob Context::prepareJob(const JobSpec& spec, const HostGraph& g) const {
    if (spec.chains <= 0) throw Error(Status::InvalidArg, "job needs at least one chain");
    if (!spec.initialSpins || !spec.seeds || !spec.runEnd || !spec.runVal || !spec.chainRunOff ||
        spec.nRuns < 1)
        throw Error(Status::InvalidArg, "job is missing spins, seeds or beta tables");
    if (spec.phases.size() != spec.iters.size())
        throw Error(Status::InvalidArg, "phases and iters differ in length");
    const int64_t n = g.n;
    const int64_t C = spec.chains;
    HostJob hj;
    hj.chains = C;
    hj.n = n;
    hj.initialSpins = spec.initialSpins;
    hj.seeds = spec.seeds;
    hj.runEnd = spec.runEnd;
    hj.runVal = spec.runVal;
    hj.nRuns = spec.nRuns;

    // Executed phases: P > 0 and at least one sweep (a P == 0 phase is
    // skipped without advancing the sweep counter, src/anneal_kernels.cu:270-272).
    uint64_t total = 0;
    int maxP = 1;
    for (size_t p = 0; p < spec.phases.size(); ++p) {
        const uint64_t P = spec.phases[p];
        if (P == 0 || spec.iters[p] == 0) continue;
        if (P > static_cast<uint64_t>(dev::kMaxGroups))
            throw Error(Status::Unsupported, "group count P=" + std::to_string(P) + " exceeds " +
                                                 std::to_string(dev::kMaxGroups));
        total += spec.iters[p];
        if (total > static_cast<uint64_t>(kMaxTotalSteps))
            throw Error(Status::StepCap, "sweeps per chain exceed the total-sweep cap");
        hj.phaseEnd.push_back(total);
        hj.phaseP.push_back(static_cast<uint32_t>(P));
        maxP = std::max<int>(maxP, static_cast<int>(P));
    }
    if (hj.phaseEnd.empty()) {   // nothing executes; keep one dummy phase for the item layout
        hj.phaseEnd.push_back(0);
        hj.phaseP.push_back(1);
    }
    hj.totalSteps = total;
    hj.maxP = maxP;

    // Per-chain table extents: a table runs from its offset to the next
    // larger distinct offset (tables are concatenated).
    hj.tableStart.assign(spec.chainRunOff, spec.chainRunOff + C);
    for (int64_t o : hj.tableStart)
        if (o < 0 || o >= spec.nRuns) throw Error(Status::InvalidArg, "chain_run_off out of range");
    std::vector<int64_t> uniq = hj.tableStart;
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    auto tableLen = [&](int64_t off) {
        const auto it = std::upper_bound(uniq.begin(), uniq.end(), off);
        return (it == uniq.end() ? spec.nRuns : *it) - off;
    };
    int64_t betaMax = 0;
    for (int64_t off : uniq) {
        const int64_t len = tableLen(off);
        if (len > INT32_MAX) throw Error(Status::Unsupported, "a beta table has more than 2^31 runs");
        for (int64_t r = off; r < off + len; ++r) {
            betaMax = std::max<int64_t>(betaMax, std::llabs(static_cast<long long>(spec.runVal[r])));
            if (r > off && spec.runEnd[r] < spec.runEnd[r - 1])
                throw Error(Status::InvalidArg, "beta run ends must be non-decreasing");
        }
        if (static_cast<uint64_t>(spec.runEnd[off + len - 1]) < total)
            throw Error(Status::InvalidArg, "a beta table ends before the last sweep");
    }
    hj.tableLen.resize(static_cast<size_t>(C));
    for (int64_t c = 0; c < C; ++c) hj.tableLen[c] = static_cast<int32_t>(tableLen(hj.tableStart[c]));
    if (spec.gateBetaMax >= 0) betaMax = spec.gateBetaMax;

    hj.bits = wordBitsFor(g.B);
    if (hj.bits == 0)
        throw Error(Status::Unsupported, "max weighted degree >= 2^62 is not supported");
    // Score arithmetic: int32 when the gate holds (src/host_common.cpp:296-312,
    // verdict at the largest beta), with 24-bit multiplies when |beta| < 2^23
    // and the field is below 2^23 (8/16-bit words); exact int64 otherwise.
    const int64_t gateB = std::max<int64_t>(spec.gateB, g.B);
    if (hj.bits == 64 || !scoreSafeInt32ForBeta(betaMax, gateB))
        hj.mode = dev::kScore64;   // 64-bit words always score in int64
    else if (betaMax < (int64_t{1} << 23) && hj.bits <= 16)
        hj.mode = dev::kScore24;
    else
        hj.mode = dev::kScore32;
    // The smallest tile (128 threads) has the smallest plan; launches pick a
    // tile that fits, and keep the chain state in HBM when none holds it.
    // The group tables alone must fit.
    const dev::LdsPlan plan = dev::makeLdsPlan(n, hj.bits, maxP, 128, hj.mode, true);
    if (static_cast<size_t>(plan.total) + dev::staticLds(hj.mode) > props_.sharedMemPerBlock)
        throw Error(Status::Unsupported, "group tables for P=" + std::to_string(maxP) + " (" +
                                             std::to_string(plan.total) +
                                             " B) exceed the LDS per workgroup");

    for (int64_t i = 0; i < C * n; ++i)
        if (spec.initialSpins[i] != 1 && spec.initialSpins[i] != -1)
            throw Error(Status::InvalidArg, "initial spins must be +1 or -1");

    // Initial cut of every chain by the host recompute (src/anneal_kernels.cu:754-764).
    hj.state.resize(static_cast<size_t>(C));
    for (int64_t c = 0; c < C; ++c) {
        const int64_t cut = anneallib::solver::cutValueRaw(g.offsets.data(), g.to.data(), g.w.data(), n,
                                                           spec.initialSpins + c * n);
        hj.state[c] = dev::ChainState{cut, cut, 0, 0, 0};
    }
    return hj;
}

std::unique_ptr<ResidentJob> Context::placeJob(const HostJob& hj, const DevGraph* g, hipStream_t stream,
                                               bool streamOrdered) {
    if (!g || g->n != hj.n) throw Error(Status::InvalidArg, "job prepared for another graph");
    const int64_t n = hj.n;
    const int64_t C = hj.chains;
    auto job = std::make_unique<ResidentJob>();
    job->graph = g;
    job->chains = C;
    job->totalSteps = hj.totalSteps;
    job->maxP = hj.maxP;
    job->hPhaseEnd = hj.phaseEnd;
    job->hPhaseP = hj.phaseP;
    job->bits = hj.bits;
    job->mode = hj.mode;

    hipCheck(hipSetDevice(device_), "hipSetDevice");
    hipMemPool_t pool = streamOrdered ? this->pool() : nullptr;
    auto alloc = [&](DevBuf& b, size_t bytes) {
        if (streamOrdered) b.allocAsync(bytes, pool, stream);
        else b.alloc(bytes);
    };
    job->wordStride = dev::wordBytes(n, job->bits);
    job->bestStride = dev::bestBytes(n);
    alloc(job->words, static_cast<size_t>(C * job->wordStride));
    alloc(job->best, static_cast<size_t>(C * job->bestStride));
    alloc(job->states, static_cast<size_t>(C) * sizeof(dev::ChainState));
    alloc(job->runEnd, static_cast<size_t>(hj.nRuns) * sizeof(int64_t));
    alloc(job->runVal, static_cast<size_t>(hj.nRuns) * sizeof(int32_t));
    alloc(job->phaseEnd, hj.phaseEnd.size() * sizeof(uint64_t));
    alloc(job->phaseP, hj.phaseP.size() * sizeof(uint32_t));
    job->footprint = static_cast<int64_t>(job->words.bytes() + job->best.bytes() + job->states.bytes() +
                                          job->runEnd.bytes() + job->runVal.bytes() +
                                          job->phaseEnd.bytes() + job->phaseP.bytes());

    hipCheck(hipMemcpyAsync(job->runEnd.get(), hj.runEnd, hj.nRuns * sizeof(int64_t),
                            hipMemcpyHostToDevice, stream), "upload beta ends");
    hipCheck(hipMemcpyAsync(job->runVal.get(), hj.runVal, hj.nRuns * sizeof(int32_t),
                            hipMemcpyHostToDevice, stream), "upload beta values");
    hipCheck(hipMemcpyAsync(job->phaseEnd.get(), hj.phaseEnd.data(), hj.phaseEnd.size() * sizeof(uint64_t),
                            hipMemcpyHostToDevice, stream), "upload phases");
    hipCheck(hipMemcpyAsync(job->phaseP.get(), hj.phaseP.data(), hj.phaseP.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, stream), "upload phases");
    hipCheck(hipMemcpyAsync(job->states.get(), hj.state.data(), C * sizeof(dev::ChainState),
                            hipMemcpyHostToDevice, stream), "upload chain state");

    // Initial words and best bits from the spins, on the device.
    DevBuf spins;
    alloc(spins, static_cast<size_t>(C * n));
    hipCheck(hipMemcpyAsync(spins.get(), hj.initialSpins, C * n, hipMemcpyHostToDevice, stream),
             "upload spins");
    std::vector<dev::InitItem> init(static_cast<size_t>(C));
    for (int64_t c = 0; c < C; ++c) {
        init[c].offsets = g->offsets.as<uint32_t>();
        init[c].edges = g->edges.as<int32_t>();
        init[c].spins = spins.as<int8_t>() + c * n;
        init[c].words = reinterpret_cast<uint32_t*>(job->words.as<uint8_t>() + c * job->wordStride);
        init[c].bestbits = reinterpret_cast<uint32_t*>(job->best.as<uint8_t>() + c * job->bestStride);
        init[c].n = static_cast<int32_t>(n);
        init[c].pad = 0;
    }
    DevBuf dInit;
    alloc(dInit, init.size() * sizeof(dev::InitItem));
    hipCheck(hipMemcpyAsync(dInit.get(), init.data(), dInit.bytes(), hipMemcpyHostToDevice, stream),
             "upload init items");
    // Chain c of a grid-y launch: gridDim.y is limited to 65535 per launch.
    for (int64_t c0 = 0; c0 < C; c0 += 65535) {
        const int cnt = static_cast<int>(std::min<int64_t>(65535, C - c0));
        hipCheck(launch::initWords(job->bits, dInit.as<dev::InitItem>() + c0, cnt, static_cast<int>(n), stream),
                 "initWords");
    }
    hipCheck(hipStreamSynchronize(stream), "admission");

    job->items.resize(static_cast<size_t>(C));
    job->stepsDone.assign(static_cast<size_t>(C), 0);
    for (int64_t c = 0; c < C; ++c) {
        dev::ChainItem& it = job->items[c];
        std::memset(&it, 0, sizeof(it));
        it.offsets = g->offsets.as<uint32_t>();
        it.edges = g->edges.as<int32_t>();
        it.words = reinterpret_cast<uint32_t*>(job->words.as<uint8_t>() + c * job->wordStride);
        it.bestbits = reinterpret_cast<uint32_t*>(job->best.as<uint8_t>() + c * job->bestStride);
        it.st = job->states.as<dev::ChainState>() + c;
        it.betaEnd = job->runEnd.as<int64_t>() + hj.tableStart[c];
        it.betaVal = job->runVal.as<int32_t>() + hj.tableStart[c];
        it.nRuns = hj.tableLen[c];
        it.phaseEnd = job->phaseEnd.as<uint64_t>();
        it.phaseP = job->phaseP.as<uint32_t>();
        it.nPhases = static_cast<int32_t>(hj.phaseEnd.size());
        it.key = hj.seeds[c];
        it.stepEnd = hj.totalSteps;
        it.n = static_cast<int32_t>(n);
        it.deltaMax = static_cast<int32_t>(std::min<int64_t>(g->B, INT32_MAX));
        it.rows16 = g->rows16.bytes() ? g->rows16.as<uint16_t>() : nullptr;
        it.m2 = static_cast<uint32_t>(g->m2);
        it.maxDeg = static_cast<int32_t>(std::min<int64_t>(g->maxDeg, INT32_MAX));
    }
    return job;
}

float Context::runSlice(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs) {
    SliceHandle h = launchSlice(jobs, knobs);
    return waitSlice(h);
}

uint64_t Context::groupKey(int bits, int mode, bool hbm, int64_t topN, bool split) {
    int cls = 0;   // size class: the largest power of two (>= 2^11) below topN
    if (split)
        for (int64_t edge = int64_t{1} << 11; edge < topN; edge <<= 1) ++cls;
    return (static_cast<uint64_t>(hbm) << 40) | (static_cast<uint64_t>(bits) << 32) |
           (static_cast<uint64_t>(mode) << 24) | static_cast<uint64_t>(cls);
}

void Context::finishGroup(LaunchGroup& g, int64_t maxN, int maxP, const LaunchKnobs& knobs) {
    int tuned = 0;
    for (const auto& kv : knobs.groupThreads)
        if (kv.first == g.key) tuned = kv.second;
    g.maxP = maxP;
    const int64_t nItems = static_cast<int64_t>(g.items.size());
    g.threads = knobs.threads > 0 ? knobs.threads
                : tuned > 0      ? tuned
                                 : chooseThreads(g.bits, g.mode, maxN, nItems, maxP, g.hbm);
    int64_t maxM2 = 0;
    for (const dev::ChainItem& it : g.items) {
        if (!it.rows16) { maxM2 = -1; break; }
        maxM2 = std::max<int64_t>(maxM2, it.m2);
    }
    shapeGroup(this, props_, g, maxN, maxM2, nItems, knobs);
}

namespace {
// Directed edges beyond which compact rows are never planned into LDS (their
// 2 B per edge alone would exceed any workgroup's LDS).
constexpr int64_t kMaxLdsRowsM2 = int64_t{1} << 17;
}  // namespace

void shapeGroup(Context* ctx, const hipDeviceProp_t& props, GroupShape& g, int64_t maxN, int64_t maxM2, int64_t nItems,
                const LaunchKnobs& knobs) {
    // LDS plan, score arithmetic and grid for one launch group.
    auto occupancy = [&](int bits, int mode, int threads, size_t lds, bool hbm, bool rich, bool rows = false) {
        return ctx ? ctx->occupancy(bits, mode, threads, lds, hbm, rich, rows)
                   : launch::maxActiveBlocks(bits, mode, threads, hbm, rich, rows, lds);
    };
    if (!dev::validThreads(g.threads))
        throw Error(Status::InvalidArg, "threads per chain must be 128, 256, 512 or 1024");
    g.plan = dev::makeLdsPlan(maxN, g.bits, g.maxP, g.threads, g.mode, g.hbm);
    if (static_cast<size_t>(g.plan.total) + dev::staticLds(g.mode) > props.sharedMemPerBlock)
        throw Error(Status::Unsupported, "combined launch needs " + std::to_string(g.plan.total) +
                                             " B of LDS per workgroup");
    int occ = std::max(1, occupancy(g.bits, g.mode, g.threads, g.plan.total, g.hbm, false));
    auto residentOf = [&](int o) {
        int64_t r = static_cast<int64_t>(o) * props.multiProcessorCount;
        if (knobs.gridCap > 0) r = std::min<int64_t>(r, knobs.gridCap);
        return std::min(r, nItems);
    };
    // 8-bit words with 24-bit scores: the table variant (one mad per score,
    // 8 KB more static LDS) replaces the per-sweep LUT when it keeps every
    // chain of the group resident that the LUT variant would.
    if (g.bits == 8 && g.mode == dev::kScore24 && !g.hbm && knobs.scoreTable >= 0) {
        const dev::LdsPlan tp = dev::makeLdsPlan(maxN, 8, g.maxP, g.threads, dev::kScoreTab);
        if (static_cast<size_t>(tp.total) + dev::staticLds(dev::kScoreTab) <= props.sharedMemPerBlock) {
            const int occTab = occupancy(8, dev::kScoreTab, g.threads, tp.total, false, false);
            if (occTab > 0 && (knobs.scoreTable > 0 || residentOf(occTab) >= residentOf(occ))) {
                g.mode = dev::kScoreTab;
                g.plan = tp;
                occ = occTab;
            }
        }
    }
    // The 24-bit score variants' 2- and 4-wave tiles have a register-rich
    // build (4 workgroups per CU instead of 6): taken when it keeps every
    // chain resident that the lean one would (and the launch has no other
    // group, Context::planGroups).
    g.rich = false;
    if ((g.mode == dev::kScoreTab || g.mode == dev::kScore24) && !g.hbm && g.threads <= 256 && !knobs.noRich) {
        const int occRich = occupancy(g.bits, g.mode, g.threads, g.plan.total, false, true);
        if (occRich > 0 && residentOf(occRich) >= residentOf(occ)) {
            g.rich = true;
            occ = occRich;
        }
    }
    // The graph's offsets and compact rows in LDS, for FLIP's row walk: taken
    // when they keep every chain resident that the plan without them would
    // (in practice launches of one or two chains per CU, whose sweeps are
    // latency-bound).
    g.rows = false;
    if (maxM2 >= 0 && maxM2 <= kMaxLdsRowsM2 && launch::rowsCapable(g.threads, g.hbm, g.rich) &&
        knobs.rowsInLds >= 0) {
        const dev::LdsPlan rp = dev::makeLdsPlan(maxN, g.bits, g.maxP, g.threads, g.mode, false, maxM2);
        if (static_cast<size_t>(rp.total) + dev::staticLds(g.mode) <= props.sharedMemPerBlock) {
            const int occRows = occupancy(g.bits, g.mode, g.threads, rp.total, false, g.rich, true);
            if (occRows > 0 && (knobs.rowsInLds > 0 || residentOf(occRows) >= residentOf(occ))) {
                g.plan = rp;
                g.rows = true;
                occ = occRows;
            }
        }
    }
    g.occupancy = occ;
    g.grid = static_cast<int>(residentOf(occ));
}

std::vector<Context::LaunchGroup> Context::planGroups(const std::vector<ResidentJob*>& jobs,
                                                      const LaunchKnobs& knobs, uint64_t slice) {
    // Group unfinished chains by kernel variant (word bits, score mode, state
    // in LDS or HBM); items sorted by descending n so the persistent
    // workgroups take the longest chains first.
    constexpr int kVariants = 2 * 4 * dev::kScoreModes;
    std::vector<LaunchGroup> base(kVariants);
    std::vector<int> maxP(kVariants, 1);
    auto gidx = [](bool hbm, int bits, int mode) {
        return ((hbm ? 4 : 0) + (bits == 8 ? 0 : bits == 16 ? 1 : bits == 32 ? 2 : 3)) * dev::kScoreModes + mode;
    };
    for (ResidentJob* j : jobs) {
        const bool hbm = stateInHbm(j->graph->n, j->bits, j->mode, j->maxP, knobs);
        const int gi = gidx(hbm, j->bits, j->mode);
        base[gi].hbm = hbm;
        base[gi].bits = j->bits;
        base[gi].mode = j->mode;
        for (size_t c = 0; c < j->items.size(); ++c) {
            if (j->stepsDone[c] >= j->totalSteps) continue;
            base[gi].items.push_back(j->items[c]);
            maxP[gi] = std::max(maxP[gi], j->maxP);
        }
    }
    // Size classes: a variant's chains split at n = 2^k (k >= 11) wherever
    // the smaller chains get a different tile, occupancy or score variant
    // from a plan sized for themselves (adjacent classes with the same
    // launch shape merge).  Small graphs in a mixed batch then run at the
    // occupancy their own LDS plan allows instead of the largest graph's.
    // Only for long launches: each class is a launch of its own, whose start
    // and end cost more than the occupancy gains on short slices (the
    // serving loop's); and a grid cap keeps one group per variant.
    const bool split = knobs.gridCap <= 0 && slice >= kClassMinSweeps;
    std::vector<LaunchGroup> out;
    for (int gi = 0; gi < kVariants; ++gi) {
        LaunchGroup& b = base[gi];
        if (b.items.empty()) continue;
        std::stable_sort(b.items.begin(), b.items.end(),
                         [](const dev::ChainItem& x, const dev::ChainItem& y) { return x.n > y.n; });
        std::vector<LaunchGroup> classes;
        size_t at = 0;
        while (at < b.items.size()) {
            const int64_t top = b.items[at].n;
            int64_t floorN = 0;   // the class holds n in (floorN, top]
            if (split)
                for (int64_t edge = int64_t{1} << 11; edge < top; edge <<= 1) floorN = edge;
            size_t stop = at;
            while (stop < b.items.size() && b.items[stop].n > floorN) ++stop;
            LaunchGroup c;
            c.hbm = b.hbm;
            c.bits = b.bits;
            c.mode = b.mode;
            c.key = groupKey(b.bits, b.mode, b.hbm, top, split);
            c.items.assign(b.items.begin() + static_cast<std::ptrdiff_t>(at),
                           b.items.begin() + static_cast<std::ptrdiff_t>(stop));
            finishGroup(c, top, maxP[gi], knobs);
            if (!classes.empty()) {
                LaunchGroup& prev = classes.back();
                if (prev.threads == c.threads && prev.mode == c.mode && prev.occupancy == c.occupancy) {
                    // Re-plan the merged class for its item count (the grid
                    // grows with it).
                    prev.items.insert(prev.items.end(), c.items.begin(), c.items.end());
                    prev.mode = b.mode;
                    finishGroup(prev, prev.items.front().n, maxP[gi], knobs);
                    at = stop;
                    continue;
                }
            }
            classes.push_back(std::move(c));
            at = stop;
        }
        for (LaunchGroup& c : classes) out.push_back(std::move(c));
    }
    // The register-rich build only for a launch of one group.  Its waves hold
    // up to 128 VGPRs, so beside three of its workgroups per CU a concurrent
    // group's 8-wave workgroup (two 80-VGPR waves per SIMD) finds no room,
    // and that group runs on the CUs left over; the lean build leaves it room
    // (the 960 chains of the 12 Gset instances in one batch: the 16-bit
    // group, which ends last, 17% sooner; the 256-graph batch 3% faster).
    if (out.size() > 1) {
        LaunchKnobs lean = knobs;
        lean.noRich = true;
        for (LaunchGroup& g : out)
            if (g.rich) finishGroup(g, g.items.front().n, g.maxP, lean);
    }
    // Largest work first: its launch starts first, and the others, on
    // concurrent lanes, fill its tail.
    auto work = [](const LaunchGroup& g) {
        int64_t w = 0;
        for (const dev::ChainItem& it : g.items) w += it.n;
        return w;
    };
    std::stable_sort(out.begin(), out.end(),
                     [&](const LaunchGroup& x, const LaunchGroup& y) { return work(x) > work(y); });
    if (out.size() > kMaxGroups)
        throw Error(Status::Unsupported, "a launch needs more than " + std::to_string(kMaxGroups) + " groups");
    if (!out.empty()) {
        const LaunchGroup& g = out.front();
        last_.threads = g.threads;
        last_.grid = g.grid;
        last_.bits = g.bits;
        last_.occupancy = g.occupancy;
        last_.mode = g.mode;
        last_.hbm = g.hbm;
        last_.rows = g.rows;
    }
    return out;
}

hipError_t Context::enqueueGroups(const std::vector<LaunchGroup>& groups,
                                  const std::vector<dev::ChainItem*>& items, uint64_t slice) {
    // Group i runs on lane i mod kLanes (lane 0 is stream_; the others fork
    // from it and join back), with its own work-counter pair.  Every lane is
    // joined even after an error, so a capture can always end.
    const int lanes = static_cast<int>(std::min<size_t>(kLanes, groups.size()));
    hipError_t err = hipSuccess;
    auto keep = [&](hipError_t e) { if (err == hipSuccess) err = e; };
    if (lanes > 1) {
        keep(hipEventRecord(fork_, stream_));
        for (int l = 1; l < lanes; ++l) keep(hipStreamWaitEvent(lanes_[l - 1], fork_, 0));
    }
    for (size_t i = 0; i < groups.size() && err == hipSuccess; ++i) {
        const LaunchGroup& g = groups[i];
        const size_t l = i % kLanes;
        keep(launch::anneal(g.bits, g.mode, g.threads, g.hbm, g.rich, g.rows, items[i], static_cast<int>(g.items.size()),
                            sched_.as<uint32_t>() + 2 * i, g.plan, g.grid, slice,
                            l == 0 ? stream_ : lanes_[l - 1]));
    }
    for (int l = 1; l < lanes; ++l) {
        keep(hipEventRecord(join_[l - 1], lanes_[l - 1]));
        keep(hipStreamWaitEvent(stream_, join_[l - 1], 0));
    }
    return err;
}

Context::SliceHandle Context::launchSlice(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs) {
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    SliceHandle h;
    const uint64_t slice = knobs.sliceSweeps > 0 ? static_cast<uint64_t>(knobs.sliceSweeps) : ~uint64_t{0};
    std::vector<LaunchGroup> groups = planGroups(jobs, knobs, slice);
    size_t nItems = 0;
    for (const LaunchGroup& g : groups) nItems += g.items.size();
    if (nItems > 0) {
        // The slot's previous slice must be done with its item array before
        // the staging copy is overwritten (normally long finished).
        ItemSlot& slot = slots_[nextSlot_];
        nextSlot_ ^= 1;
        if (slot.done) hipCheck(hipEventSynchronize(slot.done), "item slot");
        else hipCheck(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming), "hipEventCreate");
        const size_t bytes = nItems * sizeof(dev::ChainItem);
        slot.host.reserve(bytes);
        if (slot.dev.bytes() < bytes) slot.dev.alloc(slot.host.capacity());
        auto* staged = static_cast<dev::ChainItem*>(slot.host.get());
        size_t at = 0;
        for (const LaunchGroup& g : groups) {
            std::memcpy(staged + at, g.items.data(), g.items.size() * sizeof(dev::ChainItem));
            at += g.items.size();
        }
        hipCheck(hipEventCreate(&h.ev0), "hipEventCreate");
        hipCheck(hipEventCreate(&h.ev1), "hipEventCreate");
        hipCheck(hipMemcpyAsync(slot.dev.get(), staged, bytes, hipMemcpyHostToDevice, stream_), "upload items");
        // The bracket holds the kernel launches only (the item upload is
        // ordered before it on the same stream).
        hipCheck(hipEventRecord(h.ev0, stream_), "hipEventRecord");
        std::vector<dev::ChainItem*> ptrs;
        at = 0;
        for (const LaunchGroup& g : groups) {
            ptrs.push_back(slot.dev.as<dev::ChainItem>() + at);
            at += g.items.size();
        }
        hipCheck(enqueueGroups(groups, ptrs, slice), "annealSlice launch");
        hipCheck(hipEventRecord(slot.done, stream_), "hipEventRecord");
        hipCheck(hipEventRecord(h.ev1, stream_), "hipEventRecord");
        h.empty = false;
    }
    for (ResidentJob* j : jobs)
        for (auto& s : j->stepsDone) s = (j->totalSteps - s > slice) ? s + slice : j->totalSteps;
    return h;
}

Context::GraphedSlices::~GraphedSlices() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

std::unique_ptr<Context::GraphedSlices> Context::captureSlices(const std::vector<ResidentJob*>& jobs,
                                                               uint64_t sliceSweeps,
                                                               const LaunchKnobs& knobs) {
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    auto g = std::make_unique<GraphedSlices>();
    g->jobs = jobs;
    g->slice = sliceSweeps;
    std::vector<LaunchGroup> groups = planGroups(jobs, knobs, sliceSweeps);
    // Item arrays go up before the capture; the graph holds kernel nodes only.
    for (LaunchGroup& lg : groups) {
        g->itemBufs.emplace_back(lg.items.size() * sizeof(dev::ChainItem));
        hipCheck(hipMemcpy(g->itemBufs.back().get(), lg.items.data(), g->itemBufs.back().bytes(),
                           hipMemcpyHostToDevice), "upload items");
        (void)launch::maxActiveBlocks(lg.bits, lg.mode, lg.threads, lg.hbm, lg.rich, lg.rows, lg.plan.total);   // LDS attribute, outside the capture
    }
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    std::vector<dev::ChainItem*> ptrs;
    for (DevBuf& b : g->itemBufs) ptrs.push_back(b.as<dev::ChainItem>());
    const hipError_t err = enqueueGroups(groups, ptrs, sliceSweeps);
    const hipError_t endErr = hipStreamEndCapture(stream_, &g->graph);
    hipCheck(err, "annealSlice capture");
    hipCheck(endErr, "hipStreamEndCapture");
    hipCheck(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    return g;
}

float Context::replaySlices(GraphedSlices& g, int times) {
    hipCheck(hipSetDevice(device_), "hipSetDevice");
    hipEvent_t e0, e1;
    hipCheck(hipEventCreate(&e0), "hipEventCreate");
    hipCheck(hipEventCreate(&e1), "hipEventCreate");
    hipCheck(hipEventRecord(e0, stream_), "hipEventRecord");
    for (int i = 0; i < times; ++i) hipCheck(hipGraphLaunch(g.exec, stream_), "hipGraphLaunch");
    hipCheck(hipEventRecord(e1, stream_), "hipEventRecord");
    hipCheck(hipEventSynchronize(e1), "annealSlice graph");
    float ms = 0.0f;
    hipCheck(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const uint64_t adv = g.slice * static_cast<uint64_t>(times);
    for (ResidentJob* j : g.jobs)
        for (auto& s : j->stepsDone) s = (j->totalSteps - s > adv) ? s + adv : j->totalSteps;
    return ms;
}

float Context::waitSlice(SliceHandle& h) {
    float ms = 0.0f;
    if (h.ev1) {
        hipCheck(hipEventSynchronize(h.ev1), "annealSlice");
        hipCheck(hipEventElapsedTime(&ms, h.ev0, h.ev1), "hipEventElapsedTime");
        (void)hipEventDestroy(h.ev0);
        (void)hipEventDestroy(h.ev1);
        h.ev0 = h.ev1 = nullptr;
    }
    return ms;
}

void Context::autotune(const std::vector<ResidentJob*>& jobs, LaunchKnobs* knobs, float* spentMs, int keyP,
                       uint64_t budget) {
    // The groups the launch will run, planned without a forced tile.
    LaunchKnobs plain = *knobs;
    plain.threads = 0;
    plain.groupThreads.clear();
    // The launches that follow cover `budget` sweeps at most, in slices of
    // knobs->sliceSweeps: plan with that length, so the size classes match.
    uint64_t slice = knobs->sliceSweeps > 0 ? static_cast<uint64_t>(knobs->sliceSweeps) : ~uint64_t{0};
    if (budget > 0) slice = std::min(slice, budget);
    const std::vector<LaunchGroup> groups = planGroups(jobs, plain, slice);
    // Each group's trials run it alone; with other groups in the launch they
    // time the build the launch will run (planGroups: no register-rich build).
    if (groups.size() > 1) plain.noRich = true;
    knobs->groupThreads.clear();
    last_.tuned = true;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        const LaunchGroup& g = groups[gi];
        // The group's jobs: those whose unfinished chains it holds (a job's
        // chains share one graph, so they fall in one group).
        std::vector<ResidentJob*> sub;
        int64_t maxN = 0;
        uint64_t minLeft = ~uint64_t{0};
        std::unordered_set<const uint32_t*> held;
        for (const dev::ChainItem& it : g.items) held.insert(it.words);
        for (ResidentJob* j : jobs) {
            for (size_t c = 0; c < j->items.size(); ++c) {
                if (j->stepsDone[c] >= j->totalSteps) continue;
                if (held.count(j->items[c].words)) {
                    sub.push_back(j);
                    maxN = std::max<int64_t>(maxN, j->graph->n);
                    for (size_t d = 0; d < j->items.size(); ++d)
                        if (j->stepsDone[d] < j->totalSteps)
                            minLeft = std::min(minLeft, j->totalSteps - j->stepsDone[d]);
                }
                break;
            }
        }
        if (sub.empty()) continue;
        const uint64_t key = g.key * 0x9E3779B97F4A7C15ull ^
                             tuneKey(g.bits, 0, maxN, static_cast<int64_t>(g.items.size()), keyP > 0 ? keyP : g.maxP,
                                     g.hbm);
        const auto hit = tuned_.find(key);
        if (hit != tuned_.end()) {
            knobs->groupThreads.emplace_back(g.key, hit->second);
            continue;
        }
        // Each candidate runs one short slice of the group's live chains:
        // geometry does not change any chain's arithmetic, so the trial
        // sweeps are part of the result (and their time is part of the
        // caller's total), and a group's trials together use at most a
        // twenty-fifth of its sweeps.  Candidates go from the smallest tile
        // up; once one is a quarter slower than the best so far the larger
        // ones are not tried.
        const uint64_t span = budget > 0 ? std::min(budget, minLeft) : minLeft;
        const uint64_t trial = std::max<uint64_t>(1, std::min<uint64_t>(128, span / 100));
        int best = 0;
        float bestPerSweep = 0.0f;
        for (int i = 0; i < 4; ++i) {
            const int t = dev::kThreadChoices[i];
            if (!fitsLds(g.hbm ? 0 : maxN, g.bits, g.mode, g.maxP, t, g.hbm)) continue;
            LaunchKnobs k = plain;
            k.threads = t;
            k.sliceSweeps = static_cast<int64_t>(trial);
            const float ms = runSlice(sub, k);
            if (spentMs) *spentMs += ms;
            const float per = ms / static_cast<float>(trial);
            if (gi == 0) last_.trialMsPerSweep[i] = per;
            if (best == 0 || per < bestPerSweep) {
                best = t;
                bestPerSweep = per;
            } else if (per > 1.25f * bestPerSweep) {
                break;
            }
        }
        if (best == 0) continue;
        tuned_[key] = best;
        knobs->groupThreads.emplace_back(g.key, best);
    }
}

float Context::runToCompletion(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobs) {
    return runUntil(jobs, knobs, ~uint64_t{0});
}

float Context::runUntil(const std::vector<ResidentJob*>& jobs, const LaunchKnobs& knobsIn, uint64_t until) {
    float total = 0.0f;
    LaunchKnobs knobs = knobsIn;
    // The smallest sweep of a chain that still has sweeps to run before `until`.
    auto lowest = [&](uint64_t& at) {
        bool any = false;
        at = ~uint64_t{0};
        for (const ResidentJob* j : jobs) {
            const uint64_t lim = std::min(j->totalSteps, until);
            for (uint64_t s : j->stepsDone)
                if (s < lim) {
                    any = true;
                    at = std::min(at, s);
                }
        }
        return any;
    };
    auto sliceTo = [&](uint64_t at, uint64_t stop) -> int64_t {
        if (stop == ~uint64_t{0}) return knobsIn.sliceSweeps;   // 0: the rest in one launch
        const uint64_t room = stop - at;
        return static_cast<int64_t>(knobsIn.sliceSweeps > 0 ? std::min<uint64_t>(room, knobsIn.sliceSweeps)
                                                             : room);
    };
    uint64_t at = 0;
    // A call with nothing left to run keeps the previous launch's report.
    if (lowest(at)) last_ = LaunchReport{};
    if (knobs.autotune && knobs.threads == 0 && !knobs.hipGraph && lockstep(jobs)) {
        // Every chain is at the same sweep of the same phase list: tune the
        // geometry per phase (the best tile changes with the group count)
        // and cut slices at the phase ends.
        const ResidentJob* j0 = jobs.front();
        for (size_t ph = 0; ph < j0->hPhaseEnd.size() && lowest(at); ++ph) {
            const uint64_t end = std::min(j0->hPhaseEnd[ph], until);
            if (end <= at) continue;
            LaunchKnobs k = knobsIn;
            autotune(jobs, &k, &total, static_cast<int>(j0->hPhaseP[ph]), end - at);
            while (lowest(at) && at < end) {
                k.sliceSweeps = sliceTo(at, end);
                total += runSlice(jobs, k);
                if (progress_) reportProgress(jobs, total);
            }
        }
        return total;
    }
    if (knobs.autotune && knobs.threads == 0) autotune(jobs, &knobs, &total);
    if (knobs.hipGraph && knobs.sliceSweeps > 0 && until == ~uint64_t{0}) {
        // Every chain starts together here, so ceil(longest / slice) replays
        // of one captured slice finish them all.
        uint64_t longest = 0;
        for (const ResidentJob* j : jobs)
            for (uint64_t s : j->stepsDone) longest = std::max(longest, j->totalSteps - s);
        if (longest > 0) {
            auto g = captureSlices(jobs, static_cast<uint64_t>(knobs.sliceSweeps), knobs);
            const uint64_t reps = (longest + knobs.sliceSweeps - 1) / knobs.sliceSweeps;
            total += replaySlices(*g, static_cast<int>(reps));
        }
        return total;
    }
    while (lowest(at)) {
        LaunchKnobs k = knobs;
        k.sliceSweeps = sliceTo(at, until);
        total += runSlice(jobs, k);
        if (progress_) reportProgress(jobs, total);
    }
    return total;
}

bool Context::lockstep(const std::vector<ResidentJob*>& jobs) {
    if (jobs.empty() || jobs.front()->stepsDone.empty()) return false;
    const ResidentJob* j0 = jobs.front();
    const uint64_t at = j0->stepsDone.front();
    for (const ResidentJob* j : jobs) {
        if (j->hPhaseEnd != j0->hPhaseEnd || j->hPhaseP != j0->hPhaseP) return false;
        for (uint64_t s : j->stepsDone)
            if (s != at) return false;
    }
    return true;
}

void Context::collect(const ResidentJob& job, const JobSpec& spec, hipStream_t stream) {
    const int64_t C = job.chains;
    const int64_t n = job.graph->n;
    std::vector<uint8_t> hw(static_cast<size_t>(C * job.wordStride));
    std::vector<uint8_t> hb(static_cast<size_t>(C * job.bestStride));
    std::vector<dev::ChainState> hs(static_cast<size_t>(C));
    hipCheck(hipMemcpyAsync(hw.data(), job.words.get(), hw.size(), hipMemcpyDeviceToHost, stream), "download words");
    hipCheck(hipMemcpyAsync(hb.data(), job.best.get(), hb.size(), hipMemcpyDeviceToHost, stream), "download best");
    hipCheck(hipMemcpyAsync(hs.data(), job.states.get(), hs.size() * sizeof(dev::ChainState),
                            hipMemcpyDeviceToHost, stream), "download state");
    hipCheck(hipStreamSynchronize(stream), "download");
    const int bytesPerWord = job.bits / 8;
    for (int64_t c = 0; c < C; ++c) {
        const uint8_t* w = hw.data() + c * job.wordStride;
        const uint8_t* b = hb.data() + c * job.bestStride;
        if (spec.endSpins) {
            int8_t* out = spec.endSpins + c * n;
            for (int64_t v = 0; v < n; ++v) {
                // little-endian: the spin bit is the top bit of the word's last byte
                const uint8_t top = w[v * bytesPerWord + bytesPerWord - 1];
                out[v] = (top & 0x80u) ? int8_t{1} : int8_t{-1};
            }
        }
        if (spec.bestSpins) {
            int8_t* out = spec.bestSpins + c * n;
            for (int64_t v = 0; v < n; ++v) out[v] = ((b[v >> 3] >> (v & 7)) & 1u) ? int8_t{1} : int8_t{-1};
        }
        if (spec.bestVals) spec.bestVals[c] = hs[c].best;
        if (spec.finalVals) spec.finalVals[c] = hs[c].cur;
    }
}

}  // namespace rfmf
