// init_words.hpp -- device code shared by the kernels that build a chain's
// initial state: packed words and best bits from the initial spins, one
// group of 32 vertices per call.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "anneal_device.hpp"

namespace rfmf {
namespace dev {

// Words and best bits of vertices 32*grp .. 32*grp+31 of one chain:
// S[v] = sum_w w * spin[to] over v's sorted row (src/anneal_kernels.cu:251-256),
// word = spin bit | spin*S + bias; best bit = initial spin.  Returns the
// group's share of the initial cut: the weights of the edges (v, to) with
// to > v whose ends have different spins (src/max_cut.cpp:5-15), so the
// shares of all groups sum to the chain's cut.
template <int kBits>
__device__ inline int64_t initWordGroup(const uint32_t* __restrict__ offsets, const int32_t* __restrict__ edgesRaw,
                                        const int8_t* __restrict__ spins, uint32_t* __restrict__ words,
                                        uint32_t* __restrict__ bestbits, int n, int grp) {
    using F = WordFmt<kBits>;
    using Word = std::conditional_t<kBits == 64, uint64_t, uint32_t>;
    using Delta = std::conditional_t<kBits == 64, int64_t, int32_t>;
    const int2* edges = reinterpret_cast<const int2*>(edgesRaw);
    uint32_t bits = 0;
    uint32_t packed[kBits];
#pragma unroll
    for (int q = 0; q < kBits; ++q) packed[q] = 0u;
    int64_t cut = 0;
#pragma unroll 4
    for (int k = 0; k < 32; ++k) {
        const int v = 32 * grp + k;
        Word word = 0u;
        if (v < n) {
            Delta S = 0;
            const int8_t sv = spins[v];
            for (uint32_t i = offsets[v]; i < offsets[v + 1]; ++i) {
                const int2 e = edges[i];
                const int8_t su = spins[e.x];
                S += static_cast<Delta>(e.y) * static_cast<Delta>(su);
                if (e.x > v && su != sv) cut += e.y;
            }
            const bool up = sv > 0;
            const Delta delta = up ? S : -S;
            word = (up ? F::kSpin : Word{0}) | static_cast<Word>(delta + static_cast<Delta>(F::kBias));
            bits |= (up ? 1u : 0u) << k;
        }
        if constexpr (kBits == 64) {
            packed[2 * k] = static_cast<uint32_t>(word);
            packed[2 * k + 1] = static_cast<uint32_t>(word >> 32);
        } else if constexpr (kBits == 32) {
            packed[k] = word;
        } else {
            constexpr int per = 32 / kBits;
            packed[k / per] |= word << (kBits * (k % per));
        }
    }
#pragma unroll
    for (int q = 0; q < kBits; ++q) words[grp * kBits + q] = packed[q];
    bestbits[grp] = bits;
    return cut;
}

}  // namespace dev
}  // namespace rfmf
