// noalloc.hip -- rfmf_anneal_tables_noalloc: the one-shot anneal with every
// byte of device memory taken from a fixed file-scope array.
//
// The call does what rfmf_anneal_tables does (validation, int32 gate, word
// width, initial words and cuts, one persistent launch of the same
// annealSlice kernel, results copied out) without hipMalloc, hipFree, pinned
// host memory or host heap allocations on its path.  The CSR, initial spins,
// beta tables, seeds, phases, chain state, work items, work counters and
// output staging all live in `scratch`, laid out per call.  Device-side
// helper kernels build what the allocating path builds on the host (the
// interleaved edge array, the executed-phase table, the per-chain work items
// and table extents, the initial cuts, the decoded outputs), so no host-side
// array is needed either.  Set-up happens once per device, on the first call
// there: the array's device address (hipGetSymbolAddress), the device
// properties and two timing events.  Calls on one device are serialized by a
// per-device lock; each call overwrites the array's previous contents.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>

#include "../../../include/rfmf.h"
#include "../host/host_common.hpp"
#include "../host/max_cut.hpp"
#include "anneal_device.hpp"
#include "anneal_launch.hpp"
#include "capi_internal.hpp"
#include "engine.hpp"
#include "init_words.hpp"

namespace rfmf {
namespace noalloc {

// Largest scratch footprint of one call (64 MiB per device, held until the
// process ends).
constexpr size_t kScratchBytes = size_t{64} << 20;
constexpr int kMaxDevices = 16;

// Device memory of every call.  In host code this name is only a symbol
// handle: host code uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-call flags written by the device-side checks.
struct CallFlags {
    uint32_t shortTable;   // a chain's beta table ends before the last sweep
    uint32_t pad[15];
};

// Byte offsets of one call's regions in `scratch` (256-aligned).  The zeroed
// block (chain state, work counters, status) is contiguous.
struct Layout {
    size_t offsets, edges, to, w, spins, words, best, zeroBegin, states, sched, status, zeroEnd;
    size_t items, runEnd, runVal, runOff, seeds, phases, iters, phaseEnd, phaseP, bestVals, finalVals, total;
    int64_t wordStride, bestStride;
};

inline size_t up256(size_t x) { return (x + 255) / 256 * 256; }

Layout layoutFor(int64_t n, int64_t m2, int64_t C, int bits, int64_t R, int64_t np) {
    Layout L{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += up256(std::max<size_t>(bytes, 16));
        return o;
    };
    L.wordStride = dev::wordBytes(n, bits);
    L.bestStride = dev::bestBytes(n);
    L.offsets = take(static_cast<size_t>(n + 1) * 4);
    L.edges = take(static_cast<size_t>(m2) * 8);
    L.to = take(static_cast<size_t>(m2) * 4);
    L.w = take(static_cast<size_t>(m2) * 4);
    L.spins = take(static_cast<size_t>(C * n));   // initial spins, then the decoded outputs
    L.words = take(static_cast<size_t>(C * L.wordStride));
    L.best = take(static_cast<size_t>(C * L.bestStride));
    L.zeroBegin = at;
    L.states = take(static_cast<size_t>(C) * sizeof(dev::ChainState));
    L.sched = take(2 * sizeof(uint32_t));
    L.status = take(sizeof(CallFlags));
    L.zeroEnd = at;
    L.items = take(static_cast<size_t>(C) * sizeof(dev::ChainItem));
    L.runEnd = take(static_cast<size_t>(R) * 8);
    L.runVal = take(static_cast<size_t>(R) * 4);
    L.runOff = take(static_cast<size_t>(C) * 8);
    L.seeds = take(static_cast<size_t>(C) * 8);
    L.phases = take(static_cast<size_t>(np) * 8);
    L.iters = take(static_cast<size_t>(np) * 8);
    L.phaseEnd = take(static_cast<size_t>(std::max<int64_t>(np, 1)) * 8);
    L.phaseP = take(static_cast<size_t>(std::max<int64_t>(np, 1)) * 4);
    L.bestVals = take(static_cast<size_t>(C) * 8);
    L.finalVals = take(static_cast<size_t>(C) * 8);
    L.total = at;
    return L;
}

// ------------------------------------------------------------ device helpers

__global__ __launch_bounds__(256) void zeroWords(uint32_t* __restrict__ p, size_t words) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < words;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = 0u;
}

// Interleaved {to, weight} edge array from the CSR's two columns.
__global__ __launch_bounds__(256) void interleaveEdges(const int32_t* __restrict__ to, const int32_t* __restrict__ w,
                                                       int2* __restrict__ edges, int64_t m2) {
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < m2;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        edges[i] = make_int2(to[i], w[i]);
}

// Executed phases (P > 0 and at least one sweep): cumulative end sweep and
// P, or one dummy phase when none executes.  One thread.
__global__ void phaseTable(const uint64_t* __restrict__ phases, const uint64_t* __restrict__ iters, int np,
                           uint64_t* __restrict__ phaseEnd, uint32_t* __restrict__ phaseP) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint64_t total = 0;
    int k = 0;
    for (int p = 0; p < np; ++p) {
        if (phases[p] == 0 || iters[p] == 0) continue;
        total += iters[p];
        phaseEnd[k] = total;
        phaseP[k] = static_cast<uint32_t>(phases[p]);
        ++k;
    }
    if (k == 0) {
        phaseEnd[0] = 0;
        phaseP[0] = 1;
    }
}

struct ItemParams {
    const uint32_t* offsets;
    const int32_t* edges;
    unsigned char* words;
    unsigned char* best;
    dev::ChainState* states;
    const int64_t* runEnd;
    const int32_t* runVal;
    const int64_t* runOff;
    const uint64_t* seeds;
    const uint64_t* phaseEnd;
    const uint32_t* phaseP;
    dev::ChainItem* items;
    CallFlags* status;
    int64_t wordStride, bestStride;
    int64_t chains, nRuns;
    uint64_t total;
    int32_t n, nPhases, deltaMax, pad;
};

// Work item of chain c.  A chain's beta table runs from its offset to the
// next larger offset of any chain (tables are concatenated), or to the last
// run; it must reach the last sweep.
__global__ __launch_bounds__(256) void buildItems(ItemParams p) {
    const int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (c >= p.chains) return;
    const int64_t off = p.runOff[c];
    int64_t end = p.nRuns;
    for (int64_t d = 0; d < p.chains; ++d) {
        const int64_t o = p.runOff[d];
        if (o > off && o < end) end = o;
    }
    if (static_cast<uint64_t>(p.runEnd[end - 1]) < p.total) atomicOr(&p.status->shortTable, 1u);
    dev::ChainItem it;
    it.offsets = p.offsets;
    it.edges = p.edges;
    it.words = reinterpret_cast<uint32_t*>(p.words + c * p.wordStride);
    it.bestbits = reinterpret_cast<uint32_t*>(p.best + c * p.bestStride);
    it.st = p.states + c;
    it.betaEnd = p.runEnd + off;
    it.betaVal = p.runVal + off;
    it.phaseEnd = p.phaseEnd;
    it.phaseP = p.phaseP;
    it.key = p.seeds[c];
    it.stepEnd = p.total;
    it.n = p.n;
    it.nRuns = static_cast<int32_t>(end - off);
    it.nPhases = p.nPhases;
    it.deltaMax = p.deltaMax;
    it.rows16 = nullptr;
    it.m2 = 0;
    it.maxDeg = 0;
    p.items[c] = it;
}

struct InitParams {
    const uint32_t* offsets;
    const int32_t* edges;
    const int8_t* spins;
    unsigned char* words;
    unsigned char* best;
    dev::ChainState* states;
    int64_t wordStride, bestStride;
    int64_t chain0;
    int32_t n, pad;
};

// Initial words, best bits and cut of chain chain0 + blockIdx.y: each thread
// takes 32 vertices and adds its share of the cut to the chain's zeroed
// running and best cut (integer adds, so the sum is exact in any order).
template <int kBits>
__global__ __launch_bounds__(256) void initChains(InitParams p) {
    const int64_t c = p.chain0 + blockIdx.y;
    const int nGroups = (p.n + 31) >> 5;
    int64_t cut = 0;
    for (int grp = blockIdx.x * blockDim.x + threadIdx.x; grp < nGroups; grp += gridDim.x * blockDim.x)
        cut += dev::initWordGroup<kBits>(p.offsets, p.edges, p.spins + c * p.n,
                                         reinterpret_cast<uint32_t*>(p.words + c * p.wordStride),
                                         reinterpret_cast<uint32_t*>(p.best + c * p.bestStride), p.n, grp);
    if (cut != 0) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&p.states[c].cur), static_cast<unsigned long long>(cut));
        atomicAdd(reinterpret_cast<unsigned long long*>(&p.states[c].best), static_cast<unsigned long long>(cut));
    }
}

struct DecodeParams {
    const unsigned char* words;
    const unsigned char* best;
    const dev::ChainState* states;
    int8_t* out;
    int64_t* bestVals;
    int64_t* finalVals;
    int64_t wordStride, bestStride, chains;
    int32_t n, bytesPerWord;
};

// which 0: final spins from the words' spin bits (the top bit of a word's
// last byte, little-endian), and both cuts per chain; which 1: best spins
// from the best bits.
__global__ __launch_bounds__(256) void decode(DecodeParams p, int which) {
    const int64_t total = p.chains * p.n;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = i / p.n;
        const int64_t v = i - c * p.n;
        bool up;
        if (which == 0)
            up = (p.words[c * p.wordStride + v * p.bytesPerWord + p.bytesPerWord - 1] & 0x80u) != 0;
        else
            up = ((p.best[c * p.bestStride + (v >> 3)] >> (v & 7)) & 1u) != 0;
        p.out[i] = up ? int8_t{1} : int8_t{-1};
    }
    if (which == 0) {
        for (int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; c < p.chains;
             c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
            p.bestVals[c] = p.states[c].best;
            p.finalVals[c] = p.states[c].cur;
        }
    }
}

// ------------------------------------------------------------ per-device set-up

struct DeviceSlot {
    std::mutex mu;
    unsigned char* base = nullptr;   // this device's copy of `scratch`
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipDeviceProp_t props{};
};
DeviceSlot g_slots[kMaxDevices];

// With slot.mu held and `device` current.
void setUp(DeviceSlot& slot, int device) {
    if (slot.base) return;
    void* p = nullptr;
    hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    hipCheck(hipGetDeviceProperties(&slot.props, device), "hipGetDeviceProperties");
    hipCheck(hipEventCreate(&slot.ev0), "hipEventCreate");
    hipCheck(hipEventCreate(&slot.ev1), "hipEventCreate");
    slot.base = static_cast<unsigned char*>(p);
}

int gridFor(int64_t work, int64_t perBlock = 256, int cap = 4096) {
    const int64_t b = (work + perBlock - 1) / perBlock;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(b, cap)));
}

hipError_t launchInit(int bits, const InitParams& base, int64_t chains, int n) {
    const int groups = (n + 31) / 32;
    const int bx = std::max(1, (groups + 255) / 256);
    for (int64_t c0 = 0; c0 < chains; c0 += 65535) {
        InitParams p = base;
        p.chain0 = c0;
        const dim3 grid(bx, static_cast<unsigned>(std::min<int64_t>(65535, chains - c0)));
        switch (bits) {
            case 8: hipLaunchKernelGGL(initChains<8>, grid, dim3(256), 0, nullptr, p); break;
            case 16: hipLaunchKernelGGL(initChains<16>, grid, dim3(256), 0, nullptr, p); break;
            case 32: hipLaunchKernelGGL(initChains<32>, grid, dim3(256), 0, nullptr, p); break;
            case 64: hipLaunchKernelGGL(initChains<64>, grid, dim3(256), 0, nullptr, p); break;
            default: return hipErrorInvalidValue;
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace noalloc
}  // namespace rfmf

using rfmf::Error;
using rfmf::Status;

extern "C" rfmf_status rfmf_anneal_tables_noalloc(const rfmf_csr* g, int64_t max_weighted_deg,
                                                  const int8_t* initial_spins, int64_t chains,
                                                  const rfmf_beta_tables* beta, const uint64_t* phases,
                                                  const uint64_t* iters, int64_t n_phases,
                                                  const uint64_t* solver_seeds, const rfmf_launch_cfg* cfg,
                                                  int8_t* end_spins, int8_t* best_spins, float* kernel_ms,
                                                  int64_t* kernel_best_vals, int64_t* kernel_final_vals) {
    using namespace rfmf::noalloc;
    namespace dev = rfmf::dev;
    return rfmf::guarded([&] {
        rfmf::checkOneShot(g, initial_spins, chains, phases, iters, n_phases, solver_seeds);
        if (!beta || !beta->run_end || !beta->run_val || !beta->chain_run_off || beta->n_runs < 1)
            throw Error(Status::InvalidArg, "invalid beta tables");
        if (g->n == 0 || n_phases == 0) {
            rfmf::passThrough(g, initial_spins, chains, end_spins, best_spins, kernel_ms, kernel_best_vals,
                              kernel_final_vals);
            return;
        }
        const int64_t n = g->n, m2 = g->m2, C = chains, R = beta->n_runs;
        // ---- host checks: plain loops over the caller's arrays ----
        annealTotalSteps(iters, n_phases);
        if (n >= (int64_t{1} << 31) || m2 >= (int64_t{1} << 32))
            throw Error(Status::Unsupported, "uploadGraph: graph exceeds 32-bit ids/offsets");
        if (g->offsets[0] != 0 || static_cast<int64_t>(g->offsets[n]) != m2)
            throw Error(Status::InvalidArg, "uploadGraph: offsets[0] must be 0 and offsets[n] == m2");
        int64_t B = 0;
        for (int64_t v = 0; v < n; ++v) {
            if (g->offsets[v + 1] < g->offsets[v]) throw Error(Status::InvalidArg, "uploadGraph: offsets not monotone");
            int64_t s = 0;
            for (uint32_t i = g->offsets[v]; i < g->offsets[v + 1]; ++i) {
                if (g->edge_to[i] < 0 || g->edge_to[i] >= n)
                    throw Error(Status::InvalidArg, "uploadGraph: neighbour id out of range");
                s += std::llabs(static_cast<long long>(g->edge_weight[i]));
            }
            B = std::max(B, s);
        }
        uint64_t total = 0;
        int maxP = 1, nExec = 0;
        for (int64_t p = 0; p < n_phases; ++p) {
            if (phases[p] == 0 || iters[p] == 0) continue;
            if (phases[p] > static_cast<uint64_t>(dev::kMaxGroups))
                throw Error(Status::Unsupported, "group count P=" + std::to_string(phases[p]) + " exceeds " +
                                                     std::to_string(dev::kMaxGroups));
            total += iters[p];
            maxP = std::max<int>(maxP, static_cast<int>(phases[p]));
            ++nExec;
        }
        int64_t minOff = R;
        for (int64_t c = 0; c < C; ++c) {
            const int64_t o = beta->chain_run_off[c];
            if (o < 0 || o >= R) throw Error(Status::InvalidArg, "chain_run_off out of range");
            minOff = std::min(minOff, o);
        }
        // Every run from the first table start on belongs to some chain's
        // table; ends may only fall where a table starts.
        int64_t betaMax = 0;
        for (int64_t r = minOff; r < R; ++r) {
            betaMax = std::max<int64_t>(betaMax, std::llabs(static_cast<long long>(beta->run_val[r])));
            if (r > minOff && beta->run_end[r] < beta->run_end[r - 1]) {
                bool start = false;
                for (int64_t c = 0; c < C && !start; ++c) start = beta->chain_run_off[c] == r;
                if (!start) throw Error(Status::InvalidArg, "beta run ends must be non-decreasing");
            }
        }
        const int bits = rfmf::wordBitsFor(B);
        if (bits == 0) throw Error(Status::Unsupported, "max weighted degree >= 2^62 is not supported");
        const int64_t gateB = std::max<int64_t>(max_weighted_deg, B);
        int mode;
        if (bits == 64 || !scoreSafeInt32ForBeta(betaMax, gateB)) mode = dev::kScore64;
        else if (betaMax < (int64_t{1} << 23) && bits <= 16) mode = dev::kScore24;
        else mode = dev::kScore32;
        for (int64_t i = 0; i < C * n; ++i)
            if (initial_spins[i] != 1 && initial_spins[i] != -1)
                throw Error(Status::InvalidArg, "initial spins must be +1 or -1");
        const Layout L = layoutFor(n, m2, C, bits, R, n_phases);
        if (L.total > kScratchBytes)
            throw Error(Status::Unsupported,
                        "rfmf_anneal_tables_noalloc: this call needs " + std::to_string(L.total) +
                            " B of scratch; the allocation-free entry point holds " + std::to_string(kScratchBytes) +
                            " B per device (use rfmf_anneal_tables)");

        // ---- device ----
        const int device = cfg ? cfg->device : 0;
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw Error(Status::NoDevice, "no HIP device visible");
        if (device < 0 || device >= count || device >= kMaxDevices)
            throw Error(Status::NoDevice, "device ordinal " + std::to_string(device) + " out of range");
        DeviceSlot& slot = g_slots[device];
        std::lock_guard<std::mutex> lk(slot.mu);
        rfmf::hipCheck(hipSetDevice(device), "hipSetDevice");
        setUp(slot, device);
        const hipDeviceProp_t& props = slot.props;
        // The group tables alone must fit LDS (src/anneal_kernels.cu:552-558).
        const dev::LdsPlan tables = dev::makeLdsPlan(n, bits, maxP, 128, mode, true);
        if (static_cast<size_t>(tables.total) + dev::staticLds(mode) > props.sharedMemPerBlock)
            throw Error(Status::Unsupported, "group tables for P=" + std::to_string(maxP) + " (" +
                                                 std::to_string(tables.total) + " B) exceed the LDS per workgroup");
        unsigned char* const S = slot.base;
        auto at = [&](size_t off) { return S + off; };
        auto put = [&](size_t off, const void* src, size_t bytes, const char* what) {
            if (bytes) rfmf::hipCheck(hipMemcpy(at(off), src, bytes, hipMemcpyHostToDevice), what);
        };
        put(L.offsets, g->offsets, static_cast<size_t>(n + 1) * 4, "upload offsets");
        put(L.to, g->edge_to, static_cast<size_t>(m2) * 4, "upload edges");
        put(L.w, g->edge_weight, static_cast<size_t>(m2) * 4, "upload weights");
        put(L.spins, initial_spins, static_cast<size_t>(C * n), "upload spins");
        put(L.runEnd, beta->run_end, static_cast<size_t>(R) * 8, "upload beta ends");
        put(L.runVal, beta->run_val, static_cast<size_t>(R) * 4, "upload beta values");
        put(L.runOff, beta->chain_run_off, static_cast<size_t>(C) * 8, "upload table offsets");
        put(L.seeds, solver_seeds, static_cast<size_t>(C) * 8, "upload seeds");
        put(L.phases, phases, static_cast<size_t>(n_phases) * 8, "upload phases");
        put(L.iters, iters, static_cast<size_t>(n_phases) * 8, "upload iters");

        // Chain state, work counters and status start at zero (a fill kernel:
        // the previous call's contents are still there).
        const size_t zeroWordsN = (L.zeroEnd - L.zeroBegin) / 4;
        hipLaunchKernelGGL(zeroWords, dim3(gridFor(static_cast<int64_t>(zeroWordsN))), dim3(256), 0, nullptr,
                           reinterpret_cast<uint32_t*>(at(L.zeroBegin)), zeroWordsN);
        rfmf::hipCheck(hipGetLastError(), "clear scratch");
        hipLaunchKernelGGL(interleaveEdges, dim3(gridFor(m2)), dim3(256), 0, nullptr,
                           reinterpret_cast<const int32_t*>(at(L.to)), reinterpret_cast<const int32_t*>(at(L.w)),
                           reinterpret_cast<int2*>(at(L.edges)), m2);
        rfmf::hipCheck(hipGetLastError(), "interleave edges");
        hipLaunchKernelGGL(phaseTable, dim3(1), dim3(1), 0, nullptr, reinterpret_cast<const uint64_t*>(at(L.phases)),
                           reinterpret_cast<const uint64_t*>(at(L.iters)), static_cast<int>(n_phases),
                           reinterpret_cast<uint64_t*>(at(L.phaseEnd)), reinterpret_cast<uint32_t*>(at(L.phaseP)));
        rfmf::hipCheck(hipGetLastError(), "phase table");
        ItemParams ip{};
        ip.offsets = reinterpret_cast<const uint32_t*>(at(L.offsets));
        ip.edges = reinterpret_cast<const int32_t*>(at(L.edges));
        ip.words = at(L.words);
        ip.best = at(L.best);
        ip.states = reinterpret_cast<dev::ChainState*>(at(L.states));
        ip.runEnd = reinterpret_cast<const int64_t*>(at(L.runEnd));
        ip.runVal = reinterpret_cast<const int32_t*>(at(L.runVal));
        ip.runOff = reinterpret_cast<const int64_t*>(at(L.runOff));
        ip.seeds = reinterpret_cast<const uint64_t*>(at(L.seeds));
        ip.phaseEnd = reinterpret_cast<const uint64_t*>(at(L.phaseEnd));
        ip.phaseP = reinterpret_cast<const uint32_t*>(at(L.phaseP));
        ip.items = reinterpret_cast<dev::ChainItem*>(at(L.items));
        ip.status = reinterpret_cast<CallFlags*>(at(L.status));
        ip.wordStride = L.wordStride;
        ip.bestStride = L.bestStride;
        ip.chains = C;
        ip.nRuns = R;
        ip.total = total;
        ip.n = static_cast<int32_t>(n);
        ip.nPhases = std::max(1, nExec);
        ip.deltaMax = static_cast<int32_t>(std::min<int64_t>(B, INT32_MAX));
        hipLaunchKernelGGL(buildItems, dim3(gridFor(C)), dim3(256), 0, nullptr, ip);
        rfmf::hipCheck(hipGetLastError(), "work items");
        InitParams init{};
        init.offsets = ip.offsets;
        init.edges = ip.edges;
        init.spins = reinterpret_cast<const int8_t*>(at(L.spins));
        init.words = ip.words;
        init.best = ip.best;
        init.states = ip.states;
        init.wordStride = L.wordStride;
        init.bestStride = L.bestStride;
        init.n = static_cast<int32_t>(n);
        rfmf::hipCheck(launchInit(bits, init, C, static_cast<int>(n)), "initial words");
        CallFlags st{};
        rfmf::hipCheck(hipMemcpy(&st, at(L.status), sizeof(st), hipMemcpyDeviceToHost), "read status");
        if (st.shortTable) throw Error(Status::InvalidArg, "a beta table ends before the last sweep");

        // ---- the launch: the same kernel and shape rules as the allocating path ----
        rfmf::GroupShape shape;
        shape.bits = bits;
        shape.mode = mode;
        shape.maxP = maxP;
        shape.hbm = !rfmf::fitsLds(props, n, bits, mode, maxP, 128, false);
        shape.threads = rfmf::heuristicThreads(props, bits, mode, n, C, maxP, shape.hbm);
        const rfmf::LaunchKnobs knobs;
        rfmf::shapeGroup(nullptr, props, shape, n, -1, C, knobs);   // no compact rows in the scratch layout
        rfmf::hipCheck(hipEventRecord(slot.ev0, nullptr), "hipEventRecord");
        rfmf::hipCheck(rfmf::launch::anneal(bits, shape.mode, shape.threads, shape.hbm, shape.rich, false, ip.items,
                                            static_cast<int>(C), reinterpret_cast<uint32_t*>(at(L.sched)),
                                            shape.plan, shape.grid, ~uint64_t{0}, nullptr),
                       "annealSlice launch");
        rfmf::hipCheck(hipEventRecord(slot.ev1, nullptr), "hipEventRecord");
        rfmf::hipCheck(hipEventSynchronize(slot.ev1), "annealSlice");
        float ms = 0.0f;
        rfmf::hipCheck(hipEventElapsedTime(&ms, slot.ev0, slot.ev1), "hipEventElapsedTime");
        if (kernel_ms) *kernel_ms = ms;

        // ---- results ----
        DecodeParams dp{};
        dp.words = ip.words;
        dp.best = ip.best;
        dp.states = ip.states;
        dp.out = reinterpret_cast<int8_t*>(at(L.spins));
        dp.bestVals = reinterpret_cast<int64_t*>(at(L.bestVals));
        dp.finalVals = reinterpret_cast<int64_t*>(at(L.finalVals));
        dp.wordStride = L.wordStride;
        dp.bestStride = L.bestStride;
        dp.chains = C;
        dp.n = static_cast<int32_t>(n);
        dp.bytesPerWord = bits / 8;
        const int grid = gridFor(C * n);
        hipLaunchKernelGGL(decode, dim3(grid), dim3(256), 0, nullptr, dp, 0);
        rfmf::hipCheck(hipGetLastError(), "decode final spins");
        if (end_spins)
            rfmf::hipCheck(hipMemcpy(end_spins, at(L.spins), static_cast<size_t>(C * n), hipMemcpyDeviceToHost),
                           "download final spins");
        if (kernel_best_vals)
            rfmf::hipCheck(hipMemcpy(kernel_best_vals, at(L.bestVals), static_cast<size_t>(C) * 8,
                                     hipMemcpyDeviceToHost), "download best cuts");
        if (kernel_final_vals)
            rfmf::hipCheck(hipMemcpy(kernel_final_vals, at(L.finalVals), static_cast<size_t>(C) * 8,
                                     hipMemcpyDeviceToHost), "download final cuts");
        if (best_spins) {
            hipLaunchKernelGGL(decode, dim3(grid), dim3(256), 0, nullptr, dp, 1);
            rfmf::hipCheck(hipGetLastError(), "decode best spins");
            rfmf::hipCheck(hipMemcpy(best_spins, at(L.spins), static_cast<size_t>(C * n), hipMemcpyDeviceToHost),
                           "download best spins");
        }
    });
}
