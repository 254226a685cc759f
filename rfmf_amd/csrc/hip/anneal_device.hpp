// anneal_device.hpp -- device-visible data layout shared by the HIP kernels
// and the host runtime that builds their work lists.
//
// HBM layout (all chains of all graphs in one address space):
//   graph   offsets u32[n+1]; edges int2[2m] = {dense neighbour id, weight},
//           rows sorted by neighbour id (the order the reference's CSR
//           build sorts each row's edgeTo_/edgeWeight_ slice into,
//           src/immutable_graph.hpp:83-101).  Read-only, shared by every chain
//           of the graph.  offsets are padded to 16 B.
//           Graphs whose weights are all +1 or -1 and whose n is at most 32768
//           also carry compact rows u16[2m] (neighbour id | 0x8000 for weight
//           -1, padded to 16 B), which launches with LDS to spare copy into
//           LDS beside the offsets.
//   chain   words   : packed per-vertex state, one word per vertex
//                     (spin bit | delta + bias, delta = spin*S = the cut gain
//                     of flipping the vertex), padded to 16 B.
//           bestbits: 1 bit per vertex (bit v&31 of u32 v>>5), padded to 16 B.
//           ChainState (32 B): running cut, best cut, executed sweeps,
//                     improved flag.
//   schedule run-length beta table (int64 exclusive end step, int32 Q20
//           value) per chain and the phase list (cumulative executed-sweep
//           end, group count P) per chain.
//
// LDS layout (per workgroup = one chain at a time): words, bestbits, the
// per-wave group partials, the resolved winners, the Gumbel LUT, and when the
// plan has room, the graph's offsets and compact rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rfmf {
namespace dev {

struct alignas(8) ChainState {
    int64_t cur;        // incremental cut after the last executed sweep
    int64_t best;       // best incremental cut so far
    uint64_t step;      // sweeps executed (0-based id of the next sweep)
    int32_t improved;   // the last executed sweep raised best
    int32_t pad;
};
static_assert(sizeof(ChainState) == 32, "ChainState layout");

struct alignas(16) ChainItem {
    const uint32_t* offsets;   // graph CSR
    const int32_t* edges;      // interleaved {to, weight} pairs
    uint32_t* words;           // chain state words (HBM home between slices)
    uint32_t* bestbits;        // chain best-spin bitmask
    ChainState* st;
    const int64_t* betaEnd;    // run table: 0-based exclusive end step per run
    const int32_t* betaVal;    // run table: Q20 beta per run
    const uint64_t* phaseEnd;  // cumulative executed-sweep end of each P>0 phase
    const uint32_t* phaseP;    // group count of each P>0 phase
    uint64_t key;              // Philox key (solver seed)
    uint64_t stepEnd;          // run this item until st->step == stepEnd
    int32_t n;                 // vertices
    int32_t nRuns;
    int32_t nPhases;
    int32_t deltaMax;          // bound on |delta| (max weighted degree, capped at 2^31 - 1)
    const uint16_t* rows16;    // compact rows (nullptr when the graph has none)
    uint32_t m2;               // directed edges (2m)
    int32_t maxDeg;            // longest row (0: not known)
};
static_assert(sizeof(ChainItem) == 128, "ChainItem layout");

// Compact row entry: neighbour id in bits 0-14, bit 15 set for weight -1.
constexpr int64_t kCompactMaxN = 32768;
constexpr uint32_t kCompactNeg = 0x8000u;

// Word encodings: spin bit at the top of the word, delta = spin*S biased to
// stay non-negative so that 32-bit LDS atomic adds of signed changes never
// carry or borrow across neighbouring words packed in the same dword (every
// partial sum of a neighbour update is itself a valid delta, |delta| <= B, so
// the biased field stays in range at all times).
template <int kBits> struct WordFmt;
template <> struct WordFmt<8> {
    static constexpr uint32_t kSpin = 0x80u, kField = 0x7Fu;
    static constexpr int32_t kBias = 64;
    static constexpr int64_t kMaxB = 63;
};
template <> struct WordFmt<16> {
    static constexpr uint32_t kSpin = 0x8000u, kField = 0x7FFFu;
    static constexpr int32_t kBias = 16384;
    static constexpr int64_t kMaxB = 16383;
};
template <> struct WordFmt<32> {
    static constexpr uint32_t kSpin = 0x80000000u, kField = 0x7FFFFFFFu;
    static constexpr int32_t kBias = 1 << 30;
    static constexpr int64_t kMaxB = (1 << 30) - 1;
};
// Weighted degrees of 2^30 and beyond (the reference's 64-bit S,
// main.cu:51): one 64-bit word per vertex.
template <> struct WordFmt<64> {
    static constexpr uint64_t kSpin = 1ull << 63, kField = (1ull << 63) - 1;
    static constexpr int64_t kBias = int64_t{1} << 62;
    static constexpr int64_t kMaxB = (int64_t{1} << 62) - 1;
};

// Bytes of a chain's word array: n words rounded up to whole groups of 32
// vertices (the unit the initial-state kernel writes), a multiple of 16 B.
__host__ __device__ inline int64_t wordBytes(int64_t n, int bits) {
    return ((n + 31) / 32) * 4 * bits;
}
// Bytes of a chain's best-spin bitmask (1 bit per vertex, padded to 16 B).
__host__ __device__ inline int64_t bestBytes(int64_t n) {
    return ((n + 127) / 128) * 16;
}

// Largest P the kernels accept (LDS holds per-group winner records).
constexpr int kMaxGroups = 4096;
// Largest P served by the register/cross-lane argmax (above it, or when P
// does not divide 8*threads, groups are reduced with LDS atomics).
constexpr int kMaxLaneReduceP = 512;

}  // namespace dev
}  // namespace rfmf
