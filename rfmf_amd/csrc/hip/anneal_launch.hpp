// anneal_launch.hpp -- host-callable launchers for the annealing kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "anneal_device.hpp"

namespace rfmf {
namespace dev {

// Threads per workgroup (= per resident chain) the slice kernel is built
// for: 2, 4, 8 or 16 wave64.  256 is the default tile.
constexpr int kThreadChoices[4] = {128, 256, 512, 1024};
constexpr int kThreads = 256;
inline bool validThreads(int nt) { return nt == 128 || nt == 256 || nt == 512 || nt == 1024; }
// Static LDS of the slice kernel: the 16-entry biased Gumbel LUT (32-bit for
// the int32 score modes, 64-bit for the int64 one) plus 16 B of slack, and for
// kScoreTab the 2048-entry score table.
constexpr uint32_t kStaticLds = 144;
constexpr uint32_t kScoreTabLds = 2048 * 4;

// How a vertex's score = Gumbel + beta * spin*S is computed:
//   kScore32: int32 arithmetic mod 2^32 (exact whenever the int32 gate holds);
//   kScore24: the same with one v_mad_i32_i24 (|beta| < 2^23, 8/16-bit words);
//   kScore64: exact int64 scores, for schedules the int32 gate refuses
//             (the reference's two-pass path, src/anneal_kernels.cu:350-378);
//   kScoreTab: kScore24 for 8-bit words with the Gumbel quantile and its
//             7-bit fine offset read from one 2048-entry static table per
//             workgroup (8 KB of LDS), so a score is one v_mad_i32_i24; chosen
//             at launch when the table costs no resident chains.
enum ScoreMode : int { kScore32 = 0, kScore24 = 1, kScore64 = 2, kScoreTab = 3 };
constexpr int kScoreModes = 4;
inline uint32_t staticLds(int mode) { return mode == kScoreTab ? kStaticLds + kScoreTabLds : kStaticLds; }

// Byte offsets of the LDS regions of one workgroup (host-computed from the
// largest graph and the largest P of the launch).
struct LdsPlan {
    uint32_t wordsOff;    // packed words of the resident chain
    uint32_t bestOff;     // best-spin bitmask
    uint32_t partOff;     // per-wave group partials / atomic slots (u64)
    uint32_t partIdOff;   // int64 scores only: ~id beside each partial (u32)
    uint32_t pickOff;     // {winner id, pre-flip spin} per group (int32 pairs)
    uint32_t deltaOff;    // winner pre-flip delta per group (int32, int64 for 64-bit words)
    uint32_t dpartOff;    // int64 per-wave cut-delta partials
    uint32_t miscOff;     // scheduler broadcast
    uint32_t rowsOff;     // graph offsets u32[n+1] (rowsOn only)
    uint32_t rowEdgesOff; // compact rows u16[2m] (rowsOn only)
    uint32_t total;       // dynamic LDS bytes
    int32_t partStride;   // u64 entries per wave in the partial table
    int32_t slotWords;    // u64 entries of the partial/slot region
    int32_t rowsOn;       // items with compact rows copy the graph into LDS
};

// One chain's initial-state build.
struct InitItem {
    const uint32_t* offsets;
    const int32_t* edges;
    const int8_t* spins;   // n initial spins, +/-1
    uint32_t* words;
    uint32_t* bestbits;
    int32_t n;
    int32_t pad;
};

inline uint32_t alignUp(uint32_t x, uint32_t a) { return (x + a - 1) / a * a; }

// hbm: the chain state stays in HBM, so the plan has no words or best bits.
// rowsM2 >= 0: room for a graph of maxN vertices and rowsM2 directed edges in
// compact form (offsets and u16 rows), which the kernel copies into LDS for
// every item that has compact rows.
inline LdsPlan makeLdsPlan(int64_t maxN, int bits, int maxP, int threads, int mode, bool hbm = false,
                           int64_t rowsM2 = -1) {
    const int64_t rowsN = maxN;
    LdsPlan p{};
    if (hbm) maxN = 0;
    const int waves = threads / 64;
    const int laneP = maxP < kMaxLaneReduceP ? maxP : kMaxLaneReduceP;
    const int slots = (waves * laneP > maxP) ? waves * laneP : maxP;
    uint32_t off = 0;
    p.wordsOff = off; off += alignUp(static_cast<uint32_t>(wordBytes(maxN, bits)), 16);
    p.bestOff = off;  off += alignUp(static_cast<uint32_t>(bestBytes(maxN)), 16);
    p.partOff = off;  off += alignUp(static_cast<uint32_t>(slots) * 8u, 16);
    p.partIdOff = off;
    if (mode == kScore64) off += alignUp(static_cast<uint32_t>(slots) * 4u, 16);
    p.pickOff = off;  off += alignUp(static_cast<uint32_t>(maxP) * 8u, 16);
    p.deltaOff = off; off += alignUp(static_cast<uint32_t>(maxP) * (bits == 64 ? 8u : 4u), 16);
    p.dpartOff = off; off += alignUp(static_cast<uint32_t>(waves) * 8u, 16);
    p.miscOff = off;  off += 16;
    if (rowsM2 >= 0) {
        p.rowsOn = 1;
        p.rowsOff = off;     off += alignUp(static_cast<uint32_t>(rowsN + 1) * 4u, 16);
        p.rowEdgesOff = off; off += alignUp(static_cast<uint32_t>(rowsM2) * 2u, 16);
    }
    p.total = off;
    p.partStride = laneP;
    p.slotWords = slots;
    return p;
}

}  // namespace dev

namespace launch {

// Host stub of the slice kernel variant for `threads` threads per chain (and
// the register-rich and rows-in-LDS builds), or nullptr when that variant is
// not built.  Defined in anneal_slice.hpp and instantiated, one word width
// and score mode per unit, in the slice_w*.hip files.
template <int kBits, int kMode, bool kHbm>
const void* sliceKernel(int threads, bool rich, bool rows);

// Persistent slice kernel over `nItems` chain items, each advanced by at most
// `sliceSweeps` sweeps, `threads` threads per workgroup; `sched` is a 2-word
// device counter pair, zero before the launch and left zero after it.
// `hbm`: the chains' state stays in HBM (no LDS copy).  `rich`: the 24-bit
// score variants' 2- and 4-wave tiles (state in LDS) built for 5 resident
// workgroups per CU instead of 6 (more registers for the scoring loop's
// schedule); ignored elsewhere.  `rows`: the rows-in-LDS build
// (rowsCapable variants only; must equal plan.rowsOn).
hipError_t anneal(int bits, int mode, int threads, bool hbm, bool rich, bool rows, const dev::ChainItem* items,
                  int nItems, uint32_t* sched, const dev::LdsPlan& plan, int grid, uint64_t sliceSweeps,
                  hipStream_t s);

// Whether a variant carries the rows-in-LDS path (LdsPlan::rowsOn): state in
// LDS, and the register-rich build or 8- and 16-wave tiles.
inline bool rowsCapable(int threads, bool hbm, bool rich) { return !hbm && (rich || threads >= 512); }

// Resident workgroups per CU for a variant at a given dynamic LDS size.
int maxActiveBlocks(int bits, int mode, int threads, bool hbm, bool rich, bool rows, size_t lds);

// Device-to-device copy of `bytes` (a multiple of 16) with a plain kernel.
hipError_t streamCopy(const void* src, void* dst, size_t bytes, hipStream_t s);

// Initial words / best bits for `nItems` chains (maxN = largest n).
hipError_t initWords(int bits, const dev::InitItem* items, int nItems, int maxN, hipStream_t s);

}  // namespace launch
}  // namespace rfmf
