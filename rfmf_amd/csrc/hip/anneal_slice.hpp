// anneal_slice.hpp -- the CDNA4 (gfx950) slice kernel of the multi-chain
// rejection-free multiple-flip annealer.  Its variants are instantiated by
// the slice_w*.hip files, one word width and score mode each, so they
// compile in parallel.
//
// Semantics follow the reference kernel mcKernel
// (src/anneal_kernels.cu:208-491): per sweep, SELECT scores every vertex from
// the pre-sweep state (score = Gumbel + beta * spin*S, Philox4x32-10 keyed by
// the chain's solver seed, counter (sweep, v>>3)) and keeps the per-group
// argmax of group g = v mod P (ties to the smaller id); FLIP flips every
// group's winner and advances the cut exactly as sequential group-order
// flips would (the winner-winner coupling corr[]), then updates the fields.
//
// The execution is MI355X-native and differs from the reference throughout:
//   * one workgroup of 128-1024 threads (2-16 wave64) anneals one chain at a
//     time, and the chain's whole state lives in LDS for the length of a
//     slice: one packed word per vertex holding the spin bit and delta =
//     spin*S (biased to stay non-negative; 8/16/32/64 bits chosen from the
//     graph's max weighted degree) plus a best-spin bitmask;
//   * scoring is one v_mad_i32_i24 per vertex: the bias is folded either into
//     a per-sweep copy of the Gumbel LUT (LUT - bias*beta, wrapping mod 2^32;
//     the final score is exact because it fits int32 under the gate) or, for
//     8-bit words, into a static 2048-entry table of Gumbel values (the
//     scores all shift by bias*beta, which leaves the argmax unchanged); when
//     the gate fails, the same kernel scores in exact int64 and resolves
//     (score, id) pairs instead of packed 64-bit keys;
//   * Philox's block-independent half of rounds 1-3 is computed once per
//     sweep on the scalar unit;
//   * the per-group argmax is kept in registers (each lane's 8 vertex slots
//     map to fixed groups when P divides 8*threads; for int32 scores as one
//     packed key per slot, score and block number, kept with one v_max_i32,
//     in every tile with the registers for it), combined across lanes as
//     64-bit (key, ~slot) keys by a reduce-scatter on DPP and
//     v_permlane{16,32}_swap, and across waves through LDS: no atomics on the
//     hot path;
//   * the winner map for corr[] is implicit: vertex x can only win group
//     x mod P, so "x is a winner" is pick[x mod P] == x and no n-sized
//     winnerGroup array exists;
//   * FLIP updates non-winner neighbours in delta space with 32-bit LDS
//     atomics on the packed words; a winner's own word takes its flip in one
//     add and each adjacent winner's coupling in one more (or, where such
//     partial sums could leave the field, the couplings summed over the
//     winner's lanes first);
//   * the kernel is persistent over a work list of (graph, chain) items and
//     runs each item for a slice of sweeps, saving the state to HBM at the end,
//     so the same kernel serves single-graph runs, multi-graph batches and the
//     serving loop.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "anneal_device.hpp"
#include "anneal_launch.hpp"

namespace rfmf {
namespace dev {

// Q20 Gumbel quantile table coarse[i] = round(-ln(-ln((2i+1)/32)) * 2^20);
// the values must equal the reference's (src/philox.cuh:31-34) for the
// Gumbel stream to be bit-identical.
__constant__ int32_t kGumbelCoarse[16] = {
    -1303301, -903532, -648633, -438929, -249398, -68827,  109563,  290966,
    480289,   683080,  906615,  1161749, 1466888, 1858654, 2430921, 3617486};

// Largest |Gumbel| in Q20: coarse[15] + 127 (src/philox.cuh:46-48).
constexpr uint64_t kGumbelMaxMag = 3617613u;

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (Salmon et al. 2011) of the counter (sLo, sHi, b, 0) under
// the key (k0, k1): 10 rounds of
//   (c0, c1, c2, c3) <- (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0)),
// the key bumped by (W0, W1) between rounds.  Only c2 = b differs between
// the blocks of a sweep, so the products and xors that see no b are done
// once per sweep on the scalar unit (PhiloxSweep): round 1 needs one vector
// product, round 2 one, rounds 3-10 two each.
struct PhiloxSweep {
    uint32_t x1;     // sHi ^ k0: round 1 c0 = hi(M1 b) ^ x1
    uint32_t u2a;    // round 2 c0 = lo(M1 b) ^ u2a
    uint32_t u2b;    // round 2 c2 = hi(M0 c0) ^ u2b
    uint32_t u3a;    // round 3 c0 = hi(M1 c2) ^ u3a
    uint32_t k0, k1; // the key
};
__device__ __forceinline__ PhiloxSweep philoxSweep(uint32_t sLo, uint32_t sHi, uint32_t k0, uint32_t k1) {
    PhiloxSweep p;
    const uint64_t q0 = static_cast<uint64_t>(kPhiloxM0) * sLo;      // round 1, c0 side
    const uint32_t c2r1 = static_cast<uint32_t>(q0 >> 32) ^ k1;        // round 1 c2 (c3 = 0)
    const uint32_t c3r1 = static_cast<uint32_t>(q0);                   // round 1 c3
    const uint64_t q1 = static_cast<uint64_t>(kPhiloxM1) * c2r1;      // round 2, c2 side
    p.x1 = sHi ^ k0;
    p.u2a = static_cast<uint32_t>(q1 >> 32) ^ (k0 + kPhiloxW0);
    p.u2b = c3r1 ^ (k1 + kPhiloxW1);
    p.u3a = static_cast<uint32_t>(q1) ^ (k0 + 2u * kPhiloxW0);        // round 2 c1 = lo(q1)
    p.k0 = k0;
    p.k1 = k1;
    return p;
}
// A copy of a wave-uniform value in a vector register: the compiler keeps
// uniform values in SGPRs, and the kernel's many uniform values overflow them
// (spilled SGPRs live in VGPR lanes, and every use reloads them with a
// v_readlane); values a VALU instruction reads as an operand cost nothing
// extra in a VGPR.
__device__ __forceinline__ uint32_t vgprCopy(uint32_t x) {
    uint32_t v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}

// The round keys of rounds 3-10 that the per-block rounds read: (k1 + 2 W1)
// and (k0 + r W0, k1 + r W1) for r = 3..9.  kVgpr holds them in vector
// registers (the builds with registers to spare), which leaves their scalar
// registers to the sweep's other uniform values.
struct PhiloxKeys {
    uint32_t k1r2;
    uint32_t k0r[7], k1r[7];
    uint32_t m0, m1;
};
template <bool kVgpr>
__device__ __forceinline__ PhiloxKeys philoxKeys(uint32_t k0, uint32_t k1) {
    PhiloxKeys K;
    K.k1r2 = k1 + 2u * kPhiloxW1;
    K.m0 = kPhiloxM0;
    K.m1 = kPhiloxM1;
#pragma unroll
    for (uint32_t r = 3; r < 10; ++r) {
        K.k0r[r - 3] = k0 + r * kPhiloxW0;
        K.k1r[r - 3] = k1 + r * kPhiloxW1;
    }
    if constexpr (kVgpr) {
        K.k1r2 = vgprCopy(K.k1r2);
        K.m0 = vgprCopy(K.m0);
        K.m1 = vgprCopy(K.m1);
#pragma unroll
        for (int r = 0; r < 7; ++r) {
            K.k0r[r] = vgprCopy(K.k0r[r]);
            K.k1r[r] = vgprCopy(K.k1r[r]);
        }
    }
    return K;
}
__device__ __forceinline__ void philoxBlock(const PhiloxSweep& s, const PhiloxKeys& K, uint32_t b,
                                            uint32_t (&o)[4]) {
    const uint64_t p1 = static_cast<uint64_t>(K.m1) * b;         // round 1
    uint32_t c0 = static_cast<uint32_t>(p1 >> 32) ^ s.x1;
    uint32_t c1 = static_cast<uint32_t>(p1);
    const uint64_t p0 = static_cast<uint64_t>(K.m0) * c0;        // round 2
    c0 = c1 ^ s.u2a;
    uint32_t c2 = static_cast<uint32_t>(p0 >> 32) ^ s.u2b;
    uint32_t c3 = static_cast<uint32_t>(p0);
    {                                                                  // round 3
        const uint64_t a = static_cast<uint64_t>(K.m0) * c0;
        const uint64_t d = static_cast<uint64_t>(K.m1) * c2;
        c0 = static_cast<uint32_t>(d >> 32) ^ s.u3a;
        c2 = xor3(static_cast<uint32_t>(a >> 32), c3, K.k1r2);
        c1 = static_cast<uint32_t>(d);
        c3 = static_cast<uint32_t>(a);
    }
#pragma unroll
    for (uint32_t r = 3; r < 10; ++r) {
        const uint64_t a = static_cast<uint64_t>(K.m0) * c0;
        const uint64_t d = static_cast<uint64_t>(K.m1) * c2;
        const uint32_t n0 = xor3(static_cast<uint32_t>(d >> 32), c1, K.k0r[r - 3]);
        const uint32_t n2 = xor3(static_cast<uint32_t>(a >> 32), c3, K.k1r[r - 3]);
        c1 = static_cast<uint32_t>(d);
        c3 = static_cast<uint32_t>(a);
        c0 = n0;
        c2 = n2;
    }
    o[0] = c0;
    o[1] = c1;
    o[2] = c2;
    o[3] = c3;
}

// Global-memory types: pointers in address space 1 compile to global_*
// instructions with a scalar base (generic pointers would give flat_*, which
// also count against the LDS wait counter).  Native vector types, so that a
// dereference is one load (HIP's uint2/int2 are classes whose copy
// constructors would take the address back to the generic space).
template <class T>
using GlobalOf = __attribute__((address_space(1))) T;
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef int32_t i32x2 __attribute__((ext_vector_type(2)));

template <class T>
__device__ __forceinline__ const GlobalOf<T>* globalIn(const T* p) {
    return (const GlobalOf<T>*)p;
}
template <class T>
__device__ __forceinline__ GlobalOf<T>* globalIn(T* p) {
    return (GlobalOf<T>*)p;
}

// Chain-state accesses through a generic pointer that addresses HBM
// (kGlobal) or LDS.
template <bool kGlobal, class T>
__device__ __forceinline__ T loadAt(const T* p) {
    if constexpr (kGlobal) return *globalIn(p);
    else return *p;
}
template <bool kGlobal, class T>
__device__ __forceinline__ void storeAt(T* p, T v) {
    if constexpr (kGlobal) *globalIn(p) = v;
    else *p = v;
}
template <bool kGlobal>
__device__ __forceinline__ void atomicAddAt(uint32_t* p, uint32_t v) {
    if constexpr (kGlobal)
        __hip_atomic_fetch_add(globalIn(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        atomicAdd(p, v);
}

template <bool kGlobal>
__device__ __forceinline__ void atomicOrAt(uint32_t* p, uint32_t v) {
    if constexpr (kGlobal)
        __hip_atomic_fetch_or(globalIn(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        atomicOr(p, v);
}
template <bool kGlobal>
__device__ __forceinline__ void atomicAndAt(uint32_t* p, uint32_t v) {
    if constexpr (kGlobal)
        __hip_atomic_fetch_and(globalIn(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        atomicAnd(p, v);
}

template <bool kGlobal>
__device__ __forceinline__ void atomicAddAt(uint64_t* p, uint64_t v) {
    if constexpr (kGlobal)
        __hip_atomic_fetch_add(globalIn(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// A vertex's state word and its signed delta, by word width.
template <int kBits>
using WordOf = std::conditional_t<kBits == 64, uint64_t, uint32_t>;
template <int kBits>
using DeltaOf = std::conditional_t<kBits == 64, int64_t, int32_t>;

// 8 consecutive words of a vertex block, as loaded from LDS (or HBM).
template <int kBits> struct Block8 {
    static constexpr int kDw = kBits / 4;   // dwords holding 8 words
    uint32_t d[kDw];
    template <bool kGlobal = false>
    __device__ __forceinline__ void load(const uint32_t* wl, int b) {
        if constexpr (kBits == 8) {
            const u32x2 q = loadAt<kGlobal>(reinterpret_cast<const u32x2*>(wl) + b);
            d[0] = q.x; d[1] = q.y;
        } else if constexpr (kBits == 16) {
            const u32x4 q = loadAt<kGlobal>(reinterpret_cast<const u32x4*>(wl) + b);
            d[0] = q.x; d[1] = q.y; d[2] = q.z; d[3] = q.w;
        } else {
            constexpr int kQ = kBits / 16;   // u32x4 loads per block
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const u32x4 v = loadAt<kGlobal>(reinterpret_cast<const u32x4*>(wl) + kQ * b + q);
                d[4 * q] = v.x; d[4 * q + 1] = v.y; d[4 * q + 2] = v.z; d[4 * q + 3] = v.w;
            }
        }
    }
    __device__ __forceinline__ WordOf<kBits> word(int j) const {
        if constexpr (kBits == 64) {
            return static_cast<uint64_t>(d[2 * j]) | (static_cast<uint64_t>(d[2 * j + 1]) << 32);
        } else if constexpr (kBits == 32) {
            return d[j];
        } else {
            constexpr int per = 32 / kBits;
            return (d[j / per] >> (kBits * (j % per))) & ((1u << kBits) - 1u);
        }
    }
    // Biased delta field of word j.
    __device__ __forceinline__ WordOf<kBits> field(int j) const {
        if constexpr (kBits >= 32) {
            return word(j) & WordFmt<kBits>::kField;
        } else {
            constexpr int per = 32 / kBits;
            return (d[j / per] >> (kBits * (j % per))) & WordFmt<kBits>::kField;
        }
    }
    // Spin bits of the 8 words as one byte (bit j = vertex 8b+j is +1).
    // Bits above 7 of the result are not specified.
    __device__ __forceinline__ uint32_t spinByte() const {
        if constexpr (kBits == 8) {
            // Spins at bit 7 of each byte: s holds word i's at bit 8i and word
            // i+4's at 8i+4 (i < 4); shifting s by 7i moves both to bits i and
            // i+4, and no two of the OR's sources meet below bit 8.
            const uint32_t s = ((d[0] >> 7) & 0x01010101u) | ((d[1] >> 3) & 0x10101010u);
            return s | (s >> 7) | (s >> 14) | (s >> 21);
        } else if constexpr (kBits == 16) {
            // Spins at bits 15 and 31 of each dword: dword q's land at bits 2q
            // and 16 + 2q of s, and s >> 15 moves the second to 2q + 1.
            const uint32_t s = ((d[0] >> 15) & 0x00010001u) | ((d[1] >> 13) & 0x00040004u) |
                               ((d[2] >> 11) & 0x00100010u) | ((d[3] >> 9) & 0x00400040u);
            return s | (s >> 15);
        } else {
            uint32_t m = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) m |= ((word(j) >> (kBits - 1)) & 1u) << j;
            return m;
        }
    }
};

template <int kBits, bool kGlobal = false>
__device__ __forceinline__ WordOf<kBits> ldsWord(const uint32_t* wl, int x) {
    if constexpr (kBits == 64) {
        return loadAt<kGlobal>(reinterpret_cast<const uint64_t*>(wl) + x);
    } else if constexpr (kBits == 32) {
        return loadAt<kGlobal>(wl + x);
    } else {
        constexpr int per = 32 / kBits;
        return (loadAt<kGlobal>(wl + x / per) >> (kBits * (x % per))) & ((1u << kBits) - 1u);
    }
}

// Adds d (mod 2^kBits, applied without carry or borrow because the word's
// old and new values are both valid words) to word x.
template <int kBits, bool kGlobal = false>
__device__ __forceinline__ void ldsAddWord(uint32_t* wl, int x, WordOf<kBits> d) {
    if constexpr (kBits == 64) {
        atomicAddAt<kGlobal>(reinterpret_cast<uint64_t*>(wl) + x, d);
    } else {
        constexpr int per = 32 / kBits;
        const int sh = kBits * (x % per);
        atomicAddAt<kGlobal>(&wl[x / per], d << sh);
    }
}

__device__ __forceinline__ int32_t uni(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// x mod p for 0 <= x < 2^31 and p >= 2, with m = floor((2^32 - 1) / p): the
// quotient estimate mulhi(x, m) is floor(x / p) or one less, since
// x * (2^32 - p * m) < p * 2^32, so one correction makes the remainder exact.
__device__ __forceinline__ int modBy(uint32_t x, uint32_t p, uint32_t m) {
    const uint32_t r = x - __umulhi(x, m) * p;
    return static_cast<int>(r >= p ? r - p : r);
}
__device__ __forceinline__ uint32_t uni(uint32_t x) {
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int32_t>(x)));
}
__device__ __forceinline__ uint64_t uni(uint64_t x) {
    return (static_cast<uint64_t>(uni(static_cast<uint32_t>(x >> 32))) << 32) | uni(static_cast<uint32_t>(x));
}
__device__ __forceinline__ int64_t uni(int64_t x) { return static_cast<int64_t>(uni(static_cast<uint64_t>(x))); }
// Zero in a VGPR that the compiler cannot see through: an index offset that
// makes a wave-uniform load a vector load.
__device__ __forceinline__ int opaqueZero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

__device__ __forceinline__ uint64_t packKey(int32_t score, int32_t v) {
    if (v < 0) return 0ull;
    return (static_cast<uint64_t>(static_cast<uint32_t>(score) ^ 0x80000000u) << 32) |
           static_cast<uint64_t>(~static_cast<uint32_t>(v));
}

// Order-preserving unsigned image of an int64 score.  Real scores satisfy
// |score| < 2^62 (|beta| < 2^31, |delta| < 2^30), so their image is never 0,
// which therefore marks an empty slot.
__device__ __forceinline__ uint64_t biased64(int64_t s) { return static_cast<uint64_t>(s) ^ (1ull << 63); }

// (score, id) argmax with ties to the smaller id; an empty slot has v = -1
// and the type's minimum score, which no real score reaches.
template <class S>
__device__ __forceinline__ void keepBetter(S& s, int32_t& v, S s2, int32_t v2) {
    const bool take = (s2 > s) || (s2 == s && static_cast<uint32_t>(v2) < static_cast<uint32_t>(v));
    s = take ? s2 : s;
    v = take ? v2 : v;
}

// Cross-lane exchanges on the VALU (no LDS crossbar).  laneXor<d> returns
// the value of lane l ^ d for d = 1, 2 (DPP quad_perm), 4 (row_shl / row_shr
// by 4, chosen by the lane's bit 2) and 8 (row_ror by 8 within a 16-lane
// row); for 16 and 32, v_permlane16_swap / v_permlane32_swap pair each lane
// with lane l ^ 16 / l ^ 32 and return both values.
template <int kDist>
__device__ __forceinline__ int32_t rowRor(int32_t x) {
    static_assert(kDist >= 1 && kDist < 16, "row rotation distance");
    return __builtin_amdgcn_update_dpp(0, x, 0x120 + kDist, 0xF, 0xF, false);
}
template <int kDist>
__device__ __forceinline__ int32_t laneXor(int32_t x, int lane) {
    if constexpr (kDist == 1) {
        return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    } else if constexpr (kDist == 2) {
        return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    } else if constexpr (kDist == 4) {
        const int32_t up = __builtin_amdgcn_update_dpp(0, x, 0x104, 0xF, 0xF, false);     // row_shl:4 (lane + 4)
        const int32_t down = __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);   // row_shr:4 (lane - 4)
        return (lane & 4) ? down : up;
    } else {
        static_assert(kDist == 8, "lane xor distance");
        return rowRor<8>(x);
    }
}
template <int kDist>
__device__ __forceinline__ void swapPair(int32_t x, int32_t& a, int32_t& b) {
    static_assert(kDist == 16 || kDist == 32, "row/half swap distance");
    const uint32_t u = static_cast<uint32_t>(x);
    if constexpr (kDist == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        a = static_cast<int32_t>(r[0]);
        b = static_cast<int32_t>(r[1]);
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
        a = static_cast<int32_t>(r[0]);
        b = static_cast<int32_t>(r[1]);
    }
}
// The same swaps on two registers: with (a, b), lanes whose kDist bit is 0
// get (own a, partner's a) and lanes whose bit is 1 get (partner's b, own b).
template <int kDist>
__device__ __forceinline__ void swapRows(int32_t a, int32_t b, int32_t& x, int32_t& y) {
    static_assert(kDist == 16 || kDist == 32, "row/half swap distance");
    if constexpr (kDist == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(static_cast<uint32_t>(a), static_cast<uint32_t>(b), false,
                                                         false);
        x = static_cast<int32_t>(r[0]);
        y = static_cast<int32_t>(r[1]);
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(static_cast<uint32_t>(a), static_cast<uint32_t>(b), false,
                                                         false);
        x = static_cast<int32_t>(r[0]);
        y = static_cast<int32_t>(r[1]);
    }
}
__device__ __forceinline__ int32_t lo32(int64_t x) { return static_cast<int32_t>(x); }
__device__ __forceinline__ int32_t hi32(int64_t x) { return static_cast<int32_t>(x >> 32); }
__device__ __forceinline__ int64_t join64(int32_t lo, int32_t hi) {
    return (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
}

// One reduce-scatter step of the per-group argmax for a pair of slots (a, b)
// and lane distance kDist in {8, 16, 32}: lanes whose kDist bit is 0 keep slot
// a, the others slot b, each combined with its partner's copy; the result is
// left in (sa, va).  Halving the held slots per step avoids reducing every
// slot in every lane.
template <int kDist, class S>
__device__ __forceinline__ void scatterStep(S& sa, int32_t& va, const S& sb, const int32_t& vb, int lane) {
    if constexpr (kDist == 8) {
        const bool upper = (lane & 8) != 0;
        S keep = upper ? sb : sa;
        int32_t kv = upper ? vb : va;
        const S send = upper ? sa : sb;
        const int32_t sv = upper ? va : vb;
        S recv;
        if constexpr (sizeof(S) == 8)
            recv = join64(rowRor<8>(lo32(send)), rowRor<8>(hi32(send)));
        else
            recv = rowRor<8>(send);
        keepBetter(keep, kv, recv, rowRor<8>(sv));
        sa = keep;
        va = kv;
    } else {
        S x, y;
        int32_t xv, yv;
        if constexpr (sizeof(S) == 8) {
            int32_t xl, yl, xh, yh;
            swapRows<kDist>(lo32(sa), lo32(sb), xl, yl);
            swapRows<kDist>(hi32(sa), hi32(sb), xh, yh);
            x = join64(xl, xh);
            y = join64(yl, yh);
        } else {
            swapRows<kDist>(sa, sb, x, y);
        }
        swapRows<kDist>(va, vb, xv, yv);
        keepBetter(x, xv, y, yv);
        sa = x;
        va = xv;
    }
}

// One butterfly step of the per-group argmax across lanes: every lane keeps
// the better (score, id) of itself and lane l ^ kDist.  Run for kDist = Q,
// 2Q, ..., it leaves in every lane the best over the lanes that agree with it
// in every other bit (the order of a lexicographic max does not matter).
template <int kDist, class S>
__device__ __forceinline__ void foldLanes(S& s, int32_t& v, int lane) {
    if constexpr (kDist < 16) {
        S s2;
        if constexpr (sizeof(S) == 8)
            s2 = join64(laneXor<kDist>(lo32(s), lane), laneXor<kDist>(hi32(s), lane));
        else
            s2 = laneXor<kDist>(s, lane);
        keepBetter(s, v, s2, laneXor<kDist>(v, lane));
    } else {
        S sa, sb;
        int32_t va, vb;
        if constexpr (sizeof(S) == 8) {
            int32_t la, lb, ha, hb;
            swapPair<kDist>(lo32(s), la, lb);
            swapPair<kDist>(hi32(s), ha, hb);
            sa = join64(la, ha);
            sb = join64(lb, hb);
        } else {
            swapPair<kDist>(s, sa, sb);
        }
        swapPair<kDist>(v, va, vb);
        keepBetter(sa, va, sb, vb);
        s = sa;
        v = va;
    }
}

// The same reduce-scatter and butterfly steps on packed (score, ~id) keys
// (packKey): for int32 scores one unsigned 64-bit max is the whole argmax
// with ties to the smaller id, so a step is the exchange plus one compare and
// two selects, with no lane-mask logic between them.
__device__ __forceinline__ void keepMaxKey(uint64_t& a, uint64_t b) { a = b > a ? b : a; }
template <int kDist>
__device__ __forceinline__ void scatterStepKey(uint64_t& a, uint64_t b, int lane) {
    if constexpr (kDist == 8) {
        const bool upper = (lane & 8) != 0;
        uint64_t keep = upper ? b : a;
        const uint64_t send = upper ? a : b;
        const int64_t s = static_cast<int64_t>(send);
        keepMaxKey(keep, static_cast<uint64_t>(join64(rowRor<8>(lo32(s)), rowRor<8>(hi32(s)))));
        a = keep;
    } else {
        int32_t xl, yl, xh, yh;
        swapRows<kDist>(lo32(static_cast<int64_t>(a)), lo32(static_cast<int64_t>(b)), xl, yl);
        swapRows<kDist>(hi32(static_cast<int64_t>(a)), hi32(static_cast<int64_t>(b)), xh, yh);
        uint64_t x = static_cast<uint64_t>(join64(xl, xh));
        keepMaxKey(x, static_cast<uint64_t>(join64(yl, yh)));
        a = x;
    }
}
template <int kDist>
__device__ __forceinline__ void foldLanesKey(uint64_t& k, int lane) {
    const int64_t s = static_cast<int64_t>(k);
    if constexpr (kDist < 16) {
        keepMaxKey(k, static_cast<uint64_t>(join64(laneXor<kDist>(lo32(s), lane), laneXor<kDist>(hi32(s), lane))));
    } else {
        int32_t la, lb, ha, hb;
        swapPair<kDist>(lo32(s), la, lb);
        swapPair<kDist>(hi32(s), ha, hb);
        uint64_t a = static_cast<uint64_t>(join64(la, ha));
        keepMaxKey(a, static_cast<uint64_t>(join64(lb, hb)));
        k = a;
    }
}

// Sum over the 64 lanes, left in every lane: rotations inside each 16-lane
// row, then the row and half swaps (all on the VALU).
__device__ __forceinline__ int32_t waveSum32(int32_t x) {
    x += rowRor<8>(x);
    x += rowRor<4>(x);
    x += rowRor<2>(x);
    x += rowRor<1>(x);
    int32_t a, b;
    swapPair<16>(x, a, b);
    x = a + b;
    swapPair<32>(x, a, b);
    return a + b;
}

// Sum over aligned groups of L lanes (L a power of two up to 64), left in
// every lane of the group.
__device__ __forceinline__ int32_t laneGroupSum(int32_t x, int L, int lane) {
    if (L > 1) x += laneXor<1>(x, lane);
    if (L > 2) x += laneXor<2>(x, lane);
    if (L > 4) x += laneXor<4>(x, lane);
    if (L > 8) x += laneXor<8>(x, lane);
    if (L > 16) { int32_t a, b; swapPair<16>(x, a, b); x = a + b; }
    if (L > 32) { int32_t a, b; swapPair<32>(x, a, b); x = a + b; }
    return x;
}
template <int kDist>
__device__ __forceinline__ int64_t laneXor64(int64_t x, int lane) {
    return join64(laneXor<kDist>(lo32(x), lane), laneXor<kDist>(hi32(x), lane));
}
template <int kDist>
__device__ __forceinline__ int64_t swapSum64(int64_t x) {
    int32_t la, lb, ha, hb;
    swapPair<kDist>(lo32(x), la, lb);
    swapPair<kDist>(hi32(x), ha, hb);
    return join64(la, ha) + join64(lb, hb);
}
__device__ __forceinline__ int64_t laneGroupSum(int64_t x, int L, int lane) {
    if (L > 1) x += laneXor64<1>(x, lane);
    if (L > 2) x += laneXor64<2>(x, lane);
    if (L > 4) x += laneXor64<4>(x, lane);
    if (L > 8) x += laneXor64<8>(x, lane);
    if (L > 16) x = swapSum64<16>(x);
    if (L > 32) x = swapSum64<32>(x);
    return x;
}

__device__ __forceinline__ int64_t waveSum64(int64_t x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const int32_t lo = __shfl_xor(static_cast<int32_t>(x), m, 64);
        const int32_t hi = __shfl_xor(static_cast<int32_t>(x >> 32), m, 64);
        x += (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
    }
    return x;
}

// score = Gumbel + beta * delta, computed as beta * field + (LUT - bias*beta)
// + fine offset, all mod 2^32 (the field is delta + bias).
template <bool kMul24>
__device__ __forceinline__ int32_t scoreOf(int32_t beta, uint32_t field, uint32_t lutBiased, uint32_t fine) {
    const uint32_t base = lutBiased + fine;
    if constexpr (kMul24)
        return static_cast<int32_t>(static_cast<uint32_t>(__mul24(beta, static_cast<int32_t>(field))) + base);
    else
        return static_cast<int32_t>(static_cast<uint32_t>(beta) * field + base);
}

// int64 score: beta * field + (LUT - bias*beta) + fine, all mod 2^64 (the
// field is delta + bias): the exact score whenever it fits int64, as the
// reference's int64 arithmetic gives it (src/anneal_kernels.cu:350-378).
__device__ __forceinline__ int64_t scoreWide(int32_t beta, uint64_t field, int64_t lutBiased, uint32_t fine) {
    return static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(beta)) * field +
                                static_cast<uint64_t>(lutBiased) + fine);
}

// Occupancy floor per variant: 6 waves per SIMD (at most 80 VGPRs) keeps 12
// two-wave chains resident per CU in large batches; 32-bit words take 5, and
// int64 scores and HBM-resident state (graphs too large for LDS, a chain or
// two per CU) 4, where fewer registers would spill.  The score table with 8- and 16-wave tiles serves launches of one
// or two chains per CU, bound by each sweep's latency, so its registers go
// to the schedule (floor 1).  kRich: the 2- and 4-wave tiles of the 24-bit
// score variants (table and LUT, state in LDS) with a floor of 4 (128 VGPRs,
// with the Philox round keys in vector registers, which relieves the scalar
// registers), which the host picks when it keeps every chain resident that
// the lean build would; at 80 VGPRs the compiler serializes the scoring
// loop's table reads.
template <int kBits, int kMode, int NT, bool kHbm, bool kRich>
constexpr int kSliceWavesPerEu =
    kRich ? 4 : kMode == kScoreTab && NT >= 512 ? 1 : (kMode == kScore64 || kHbm) ? 4 : kBits == 32 ? 5 : 6;

// kHbm: the chain's words and best bits stay in HBM and are updated in place
// (graphs whose state does not fit LDS).  FLIP's global atomics complete in
// L2, so after the sweep's last barrier each wave invalidates its L1 (an
// agent-scope acquire) before the next sweep reads the words.
// The rows-in-LDS builds hold one chain per CU (the graph's rows take the
// LDS), so their floor is the chain's own waves per SIMD: the register budget
// goes to the schedule (ER n=800 with 64 chains 5% faster than at a floor
// of 5).
template <int NT>
constexpr int kRowsWavesPerEu = NT >= 512 ? NT / 256 : 1;

// kRows: the build that copies the graph's offsets and compact rows into LDS
// for items that have them (LdsPlan::rowsOn), for launches of about one chain
// per CU; only the register-rich 2- and 4-wave tiles and the 8- and 16-wave
// tiles have it (launch::rowsCapable), and the builds without it are the same
// code as before the path existed.
template <int kBits, int kMode, int NT, bool kHbm, bool kRich = false, bool kRows = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(kRows ? kRowsWavesPerEu<NT> : kSliceWavesPerEu<kBits, kMode, NT, kHbm, kRich>)))
void annealSlice(const ChainItem* __restrict__ items, int nItems, uint32_t* __restrict__ sched, LdsPlan plan,
                 uint64_t sliceSweeps) {
    constexpr int NW = NT / 64;
    constexpr int kLgNT = NT == 128 ? 7 : NT == 256 ? 8 : NT == 512 ? 9 : 10;
    static_assert((1 << kLgNT) == NT, "threads per chain: 128, 256, 512 or 1024");
    constexpr bool kTab = kMode == kScoreTab;
    constexpr bool kMul24 = kMode == kScore24 || kTab;
    constexpr bool kWide = kMode == kScore64;
    // Builds with at least 128 VGPRs per lane (4 waves per SIMD or fewer).
    constexpr bool kRoomy = kRows || kSliceWavesPerEu<kBits, kMode, NT, kHbm, kRich> <= 4;
    // Packed per-lane argmax keys (int32 scores); the other 8- and 16-wave
    // builds keep the (score, block) loop, which spills nothing.
    constexpr bool kPackable = !kWide && (NT <= 256 || kRoomy);
    // The per-group argmax folds packed (score, ~id) keys for int32 scores
    // (fewer dependent steps than (score, id) pairs).
    constexpr bool kKeyFold = !kWide;
    // The register-rich and rows-in-LDS builds hold the scoring loop's
    // uniform operands (Philox multipliers, round keys and per-sweep terms)
    // in vector registers, which relieves the scalar registers.
    constexpr bool kVgprUniform = kRich || kRows;
    // One straight-line key fold per group count where the register budget
    // allows it.
    constexpr bool kFoldPerP = kRoomy;
    // The register-rich builds without rows in LDS let the best-spin snapshot
    // wait for the end of a run of improving sweeps (below): the 1024-chain
    // torus gains 2%, the 80-chain G81 1%.  The rows-in-LDS builds' latency-
    // bound sweeps (one chain per CU) lost 1-2% to it, and the lean builds'
    // batches 1%, so they pack the state before the next sweep flips.
    constexpr bool kDeferBest = kRich && !kRows;
    static_assert(!kRows || (!kHbm && (kRich || NT >= 512)),
                  "rows in LDS: state in LDS, register-rich 2- and 4-wave tiles or 8- and 16-wave tiles");
    static_assert(!kTab || kBits == 8, "the score table needs 8-bit words (field < 128)");
    using Score = std::conditional_t<kWide, int64_t, int32_t>;
    constexpr Score kNoScore = kWide ? static_cast<Score>(INT64_MIN) : static_cast<Score>(INT32_MIN);
    using F = WordFmt<kBits>;
    using Word = WordOf<kBits>;
    using Delta = DeltaOf<kBits>;
    static_assert(!kHbm || !kTab, "HBM-resident state runs the LUT variants");
    static_assert(!kRich || ((kTab || kMode == kScore24) && NT <= 256 && !kHbm),
                  "register-rich builds: 24-bit scores with state in LDS, 2- and 4-wave tiles");
    static_assert(kBits < 64 || kWide, "64-bit words score in int64");
    extern __shared__ __align__(16) unsigned char lds[];
    uint32_t* const ldsWords = reinterpret_cast<uint32_t*>(lds + plan.wordsOff);
    uint32_t* const ldsBest = reinterpret_cast<uint32_t*>(lds + plan.bestOff);
    uint64_t* part = reinterpret_cast<uint64_t*>(lds + plan.partOff);
    uint32_t* partId = reinterpret_cast<uint32_t*>(lds + plan.partIdOff);   // int64 scores only
    // Winner record of group g: {id (-1: none), pre-flip spin}, one 8-byte
    // read in FLIP.
    i32x2* pks = reinterpret_cast<i32x2*>(lds + plan.pickOff);
    Delta* pd = reinterpret_cast<Delta*>(lds + plan.deltaOff);
    // Static LDS, so the table's address folds into the reads' offset field:
    // 16 x u32 (int32 scores) or 16 x int64 (int64 scores).
    __shared__ __align__(16) uint32_t lutRaw[32];
    uint32_t* lutb = lutRaw;
    int64_t* lutw = reinterpret_cast<int64_t*>(lutRaw);
    // kScoreTab: gtab[r] = LUT[r & 15] + (r >> 4 & 127) for the 11-bit field
    // r, independent of beta.  A score is then beta * field + gtab[r], the
    // true score plus bias * beta: the same shift for every vertex of a sweep,
    // so the argmax is unchanged, and |beta * field| < 2^23 * 128 keeps it
    // inside int32.  Filled once per launch, before the first barrier.
    __shared__ int32_t gtab[kTab ? 2048 : 1];
    if constexpr (kTab) {
        for (int i = threadIdx.x; i < 2048; i += NT) gtab[i] = kGumbelCoarse[i & 15] + ((i >> 4) & 0x7F);
    }
    // Lane t < 16 keeps LUT entry t in a register, so the per-sweep rebuild
    // reads no memory.
    const int32_t lutCoarse = !kTab && threadIdx.x < 16 ? kGumbelCoarse[threadIdx.x] : 0;
    auto setLut = [&](int32_t b) {
        if constexpr (kTab) return;
        const int t = threadIdx.x;
        if (t < 16) {
            if constexpr (kWide)
                lutw[t] = static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(lutCoarse)) -
                                               static_cast<uint64_t>(F::kBias) *
                                                   static_cast<uint64_t>(static_cast<int64_t>(b)));
            else
                lutb[t] = static_cast<uint32_t>(lutCoarse) - static_cast<uint32_t>(F::kBias) * static_cast<uint32_t>(b);
        }
    };
    int64_t* dpart = reinterpret_cast<int64_t*>(lds + plan.dpartOff);
    int32_t* misc = reinterpret_cast<int32_t*>(lds + plan.miscOff);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform per wave (an SGPR)
    const int pstride = plan.partStride;

    for (;;) {
        if (tid == 0) misc[0] = static_cast<int32_t>(atomicAdd(&sched[0], 1u));
        __syncthreads();
        const int itemIdx = __builtin_amdgcn_readfirstlane(misc[0]);
        if (itemIdx >= nItems) break;

        // Item fields and every per-sweep scalar are wave-uniform; they are
        // pinned to SGPRs with readfirstlane (the loads go through generic
        // pointers, which the compiler would otherwise keep in VGPRs).
        const ChainItem& it = items[itemIdx];
        const int n = uni(it.n);
        const GlobalOf<uint32_t>* __restrict__ offsets = globalIn(it.offsets);
        const GlobalOf<i32x2>* __restrict__ edges = globalIn(reinterpret_cast<const i32x2*>(it.edges));
        GlobalOf<ChainState>* st = globalIn(it.st);
        const GlobalOf<int64_t>* betaEnd = globalIn(it.betaEnd);
        const GlobalOf<int32_t>* betaVal = globalIn(it.betaVal);
        const GlobalOf<uint64_t>* phaseEnd = globalIn(it.phaseEnd);
        const GlobalOf<uint32_t>* phaseP = globalIn(it.phaseP);
        const int nRuns = uni(it.nRuns);
        const int nPhases = uni(it.nPhases);
        const uint32_t k0 = uni(static_cast<uint32_t>(it.key));
        const uint32_t k1 = uni(static_cast<uint32_t>(it.key >> 32));
        const PhiloxKeys pkeys = philoxKeys<kVgprUniform>(k0, k1);

        // ---- the chain's state: copied into LDS, or used in place ----
        const int wq = static_cast<int>(wordBytes(n, kBits) / 16);
        const int bq = static_cast<int>(bestBytes(n) / 16);
        uint32_t* wl;
        uint32_t* bl;
        if constexpr (kHbm) {
            wl = it.words;
            bl = it.bestbits;
        } else {
            wl = ldsWords;
            bl = ldsBest;
            const GlobalOf<u32x4>* gw = globalIn(reinterpret_cast<const u32x4*>(it.words));
            u32x4* lw = reinterpret_cast<u32x4*>(wl);
            for (int i = tid; i < wq; i += NT) lw[i] = gw[i];
            const GlobalOf<u32x4>* gb = globalIn(reinterpret_cast<const u32x4*>(it.bestbits));
            u32x4* lb = reinterpret_cast<u32x4*>(bl);
            for (int i = tid; i < bq; i += NT) lb[i] = gb[i];
        }
        uint8_t* const bl8 = reinterpret_cast<uint8_t*>(bl);
        // ---- the graph's rows: copied into LDS when the plan has room and
        // the graph has compact rows (FLIP's row walk is then a chain of LDS
        // reads instead of dependent L2 loads) ----
        const uint32_t* const rowOffL = reinterpret_cast<const uint32_t*>(lds + plan.rowsOff);
        const uint16_t* const rowEdgL = reinterpret_cast<const uint16_t*>(lds + plan.rowEdgesOff);
        const bool rowsInLds =
            kRows && plan.rowsOn != 0 && uni(static_cast<int32_t>(it.rows16 != nullptr)) != 0;
        if (rowsInLds) {
            const int rq = (n + 4) >> 2;   // u32x4 pieces of offsets[0..n] (padded to 16 B)
            const int eq = static_cast<int>((uni(it.m2) + 7u) >> 3);   // u32x4 pieces of the u16 rows
            const GlobalOf<u32x4>* go = globalIn(reinterpret_cast<const u32x4*>(it.offsets));
            u32x4* lo = reinterpret_cast<u32x4*>(lds + plan.rowsOff);
            for (int i = tid; i < rq; i += NT) lo[i] = go[i];
            const GlobalOf<u32x4>* ge = globalIn(reinterpret_cast<const u32x4*>(it.rows16));
            u32x4* le = reinterpret_cast<u32x4*>(lds + plan.rowEdgesOff);
            for (int i = tid; i < eq; i += NT) le[i] = ge[i];
        }
        int64_t cur = uni(st->cur);
        int64_t best = uni(st->best);
        uint64_t step = uni(st->step);
        int improved = uni(st->improved);
        const uint64_t itemEnd = uni(it.stepEnd);
        const uint64_t remaining = itemEnd > step ? itemEnd - step : 0;
        const uint64_t stepEnd = step + (remaining < sliceSweeps ? remaining : sliceSweeps);
        for (int g = tid; g < plan.slotWords; g += NT) {
            part[g] = 0ull;
            if constexpr (kWide) partId[g] = 0u;
        }

        // Phase and beta-run cursors for the slice's first sweep.
        int ph = 0;
        while (ph < nPhases - 1 && uni(phaseEnd[ph]) <= step) ++ph;
        uint64_t phaseStop = uni(phaseEnd[ph]);
        int P = static_cast<int>(uni(phaseP[ph]));
        int run = 0;
        {
            int lo = 0, hi = nRuns - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (static_cast<uint64_t>(uni(betaEnd[mid])) > step) hi = mid; else lo = mid + 1;
            }
            run = lo;
        }
        uint64_t runStop = static_cast<uint64_t>(uni(betaEnd[run]));
        int32_t beta = uni(betaVal[run]);
        setLut(beta);
        __syncthreads();

        const int nFull = n >> 3;
        const int nBlk = (n + 7) >> 3;
        // Waves that own at least one block (all of them unless the chain has
        // fewer blocks than threads).  The others skip the per-group
        // reduction: their partial entries stay zero (empty).
        const int busyWaves = nBlk >= NT ? NW : (nBlk + 63) >> 6;
        // Blocks per lane (K) and the key bits that number them.
        const int lanesK = (nBlk + NT - 1) / NT;
        const int lanesBits = lanesK <= 1 ? 0 : 32 - __clz(lanesK - 1);
        const int lanesFull = nFull / NT;   // full blocks every lane scores
        const uint64_t deltaMax = static_cast<uint64_t>(uni(it.deltaMax));
        // log2 of the lanes that walk the longest row two edges at a time
        // (ceil(maxDeg / 2) rounded up to a power of two; 64 when unknown).
        const int rowLanesLg = [&]() {
            const int d = uni(it.maxDeg);
            if (d <= 0) return 6;
            const int half = (d + 1) >> 1;
            const int lg = half <= 1 ? 0 : 32 - __clz(half - 1);
            return lg > 6 ? 6 : lg;
        }();
        // Best-spin snapshot: the spins of the words now into bl.  Each thread
        // packs its own blocks (bytes of bl).  The last block's bits past n
        // are zero without a mask (initWordGroup writes the words past n as
        // zero, and no winner or row entry names them).  The 80-VGPR builds
        // apply one anyway, since without it their registers spill to scratch,
        // and so do the deferring builds, whose schedule it suits (the
        // 1024-chain torus 1.6% faster with it than without).
        auto snapshot = [&]() {
            for (int b = tid; b < nBlk; b += NT) {
                Block8<kBits> blk;
                blk.template load<kHbm>(wl, b);
                uint32_t m = blk.spinByte();
                if constexpr (!kRoomy || kDeferBest)
                    if (8 * b + 8 > n) m &= (1u << (n - 8 * b)) - 1u;
                storeAt<kHbm>(bl8 + b, static_cast<uint8_t>(m));
            }
        };

        // kDeferBest: a sweep that raises best leaves the best state in the
        // words; it is packed only when the next sweep fails to raise best
        // again, from the words then, with that sweep's winners set back to
        // their pre-sweep spins (their records stay in pks until the next
        // RESOLVE): a run of improving sweeps (a descent) costs one pass.
        // `pending`: the state before the last sweep raised best and is not in
        // bl.  Otherwise the state is packed before the next sweep flips.  At a
        // slice's start bl already holds the best state.
        auto snapshotBeforeLast = [&]() {
            snapshot();
            __syncthreads();   // every byte of bl written
            for (int g = tid; g < P; g += NT) {
                const i32x2 rec = pks[g];
                if (rec.x >= 0) {
                    const uint32_t bit = 1u << (rec.x & 31);
                    if (rec.y > 0)
                        atomicOrAt<kHbm>(bl + (rec.x >> 5), bit);
                    else
                        atomicAndAt<kHbm>(bl + (rec.x >> 5), ~bit);
                }
            }
        };
        const uint64_t stepStart = step;
        bool pending = false;

        while (step < stepEnd) {
            if constexpr (kDeferBest) {
                // P is still the last sweep's here (the phase advances below).
                if (pending && !improved) snapshotBeforeLast();
                pending = improved != 0 && step != stepStart;
            } else {
                if (improved) snapshot();
            }
            if (step >= phaseStop && ph < nPhases - 1) {
                while (ph < nPhases - 1 && uni(phaseEnd[ph]) <= step) ++ph;
                phaseStop = uni(phaseEnd[ph]);
                P = static_cast<int>(uni(phaseP[ph]));
            }
            // The next beta run, fetched now so that its latency hides behind
            // SELECT and RESOLVE (it is consumed after RESOLVE).  Vector loads:
            // a scalar load shares its wait counter with LDS accesses, so
            // SELECT's first LDS wait would also wait for it.  The lean
            // 16-bit-word builds (80-VGPR floor) would spill the three VGPRs
            // and take scalar loads.
            int64_t nextEnd = static_cast<int64_t>(runStop);
            int32_t nextVal = beta;
            if (run < nRuns - 1) {
                if constexpr (!(kBits == 16 && !kRich && !kRows)) {
                    const int z = opaqueZero();
                    nextEnd = betaEnd[run + 1 + z];
                    nextVal = betaVal[run + 1 + z];
                } else {
                    nextEnd = uni(static_cast<int64_t>(betaEnd[run + 1]));
                    nextVal = uni(static_cast<int32_t>(betaVal[run + 1]));
                }
            }
            // The group map without integer division unless P is not a power
            // of two (a mask, else a reciprocal), and the register-argmax
            // condition (a power of two divides 8 * threads iff it does not
            // exceed it).
            const bool pow2 = (P & (P - 1)) == 0;
            const bool laneReduce = pow2 && P <= kMaxLaneReduceP && P <= 8 * NT;
            auto recipOfP = [&]() { return pow2 ? 0u : 0xFFFFFFFFu / static_cast<uint32_t>(P); };
            auto groupOf = [&](int x, uint32_t pRecip) {
                return pow2 ? (x & (P - 1)) : modBy(static_cast<uint32_t>(x), static_cast<uint32_t>(P), pRecip);
            };
            const uint64_t s1 = step + 1;   // 1-based sweep id, the Philox counter
            const uint32_t sLo = static_cast<uint32_t>(s1);
            const uint32_t sHi = static_cast<uint32_t>(s1 >> 32);
            PhiloxSweep phx = philoxSweep(sLo, sHi, k0, k1);
            if constexpr (kVgprUniform) {
                phx.x1 = vgprCopy(phx.x1);
                phx.u2a = vgprCopy(phx.u2a);
                phx.u2b = vgprCopy(phx.u2b);
                phx.u3a = vgprCopy(phx.u3a);
            }
            // Packed per-lane keys when every score of the sweep satisfies
            // |score| <= |beta| * deltaMax + kGumbelMaxMag < 2^(31 - lanesBits)
            // (the (score, block) loop otherwise).
            const uint64_t absBeta = static_cast<uint64_t>(beta < 0 ? -static_cast<int64_t>(beta) : beta);
            const int packShift = kPackable &&
                                          absBeta * deltaMax + kGumbelMaxMag < (uint64_t{1} << (31 - lanesBits))
                                      ? lanesBits
                                      : -1;

            // ---------------- SELECT ----------------
            // Per-wave partial of a group: a packed (score, ~id) key for int32
            // scores, a biased score plus ~id beside it for int64 scores.
            auto putPartial = [&](int idx, Score sc, int32_t v) {
                if constexpr (kWide) {
                    part[idx] = v >= 0 ? biased64(sc) : 0ull;
                    partId[idx] = v >= 0 ? ~static_cast<uint32_t>(v) : 0u;
                } else {
                    part[idx] = packKey(sc, v);
                }
            };
            // int64 scores with LDS slots take two passes (the max score, then
            // the smallest id holding it); this flag marks the second.
            bool idPass = false;
            // One vertex's score (int32 variants): the true score, or for the
            // table variant the true score plus bias * beta.
            auto score32 = [&](const Block8<kBits>& blk, int j, uint32_t r) {
                if constexpr (kTab)
                    return __mul24(beta, static_cast<int32_t>(blk.field(j))) + gtab[r & 0x7FFu];
                else
                    return scoreOf<kMul24>(beta, static_cast<uint32_t>(blk.field(j)), lutb[r & 15u], (r >> 4) & 0x7Fu);
            };
            // Per-lane argmax with packed keys: key = score * 2^packShift +
            // (K - 1 - k) for the lane's k-th block, so one max keeps the
            // best score and, among equal scores, the earlier block (the
            // smaller vertex).  Exact while |score| < 2^(31 - packShift),
            // which the per-sweep bound guarantees.
            auto scorePackedLanes = [&](int32_t (&acc)[8]) {
#pragma unroll
                for (int j = 0; j < 8; ++j) acc[j] = INT32_MIN;
                const int cb = kPackable ? packShift : 0;
                const uint32_t biasBeta = kTab ? static_cast<uint32_t>(F::kBias) * static_cast<uint32_t>(beta) : 0u;
                const uint32_t c0 = static_cast<uint32_t>(lanesK - 1) - (biasBeta << cb);
                auto scoreBlockPacked = [&](int b, uint32_t c, bool tail) {
                    Block8<kBits> blk;
                    blk.template load<kHbm>(wl, b);
                    uint32_t o[4];
                    philoxBlock(phx, pkeys, static_cast<uint32_t>(b), o);
                    // The byte offsets of both fields of an output word come
                    // from one shifted copy: (w << 2) & mask and its high half
                    // & mask (a word-select operand).  The copy is opaque to
                    // the compiler, which would otherwise fold it back into a
                    // shift and a mask per field.
                    uint32_t o4[4];
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        o4[w] = o[w] << 2;
                        asm("" : "+v"(o4[w]));
                    }
                    if constexpr (kRich) {
                        // The register-rich build issues the block's eight
                        // table (or LUT) reads together, ahead of every use
                        // (with few waves per SIMD nothing else hides their
                        // latency).
                        uint32_t gv[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t off = ((j & 1) ? (o4[j >> 1] >> 16) : o4[j >> 1]) & (kTab ? 0x1FFCu : 0x3Cu);
                            const unsigned char* base = kTab ? reinterpret_cast<const unsigned char*>(gtab)
                                                             : reinterpret_cast<const unsigned char*>(lutb);
                            gv[j] = *reinterpret_cast<const uint32_t*>(base + off);
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            if (tail && 8 * b + j >= n) continue;
                            uint32_t t;
                            if constexpr (kTab) {
                                t = static_cast<uint32_t>(__mul24(beta, static_cast<int32_t>(blk.field(j))) +
                                                          static_cast<int32_t>(gv[j]));
                            } else {
                                const uint32_t r = ((j & 1) ? (o[j >> 1] >> 16) : o[j >> 1]);
                                t = static_cast<uint32_t>(
                                    scoreOf<kMul24>(beta, static_cast<uint32_t>(blk.field(j)), gv[j], (r >> 4) & 0x7Fu));
                            }
                            acc[j] = max(acc[j], static_cast<int32_t>((t << cb) + c));
                        }
                        return;
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (tail && 8 * b + j >= n) continue;
                        uint32_t t;
                        if constexpr (kTab) {   // table entry r & 0x7FF
                            const uint32_t off = ((j & 1) ? (o4[j >> 1] >> 16) : o4[j >> 1]) & 0x1FFCu;
                            const int32_t g = *reinterpret_cast<const int32_t*>(
                                reinterpret_cast<const unsigned char*>(gtab) + off);
                            t = static_cast<uint32_t>(__mul24(beta, static_cast<int32_t>(blk.field(j))) + g);
                        } else {                // LUT entry r & 15
                            const uint32_t off = ((j & 1) ? (o4[j >> 1] >> 16) : o4[j >> 1]) & 0x3Cu;
                            const uint32_t lut = *reinterpret_cast<const uint32_t*>(
                                reinterpret_cast<const unsigned char*>(lutb) + off);
                            const uint32_t r = ((j & 1) ? (o[j >> 1] >> 16) : o[j >> 1]);
                            t = static_cast<uint32_t>(
                                scoreOf<kMul24>(beta, static_cast<uint32_t>(blk.field(j)), lut, (r >> 4) & 0x7Fu));
                        }
                        acc[j] = max(acc[j], static_cast<int32_t>((t << cb) + c));
                    }
                };
                // A key is one v_lshl_add_u32 (score, shift, c = K - 1 - k):
                // the shift is in an SGPR and c in a VGPR, since one VALU
                // instruction reads at most one SGPR.  Every lane runs the
                // nFull / NT full rounds (a uniform loop); the lanes below
                // nFull mod NT score one more full block, and lane nFull mod NT
                // the partial one (block nFull, the same position
                // k = lanesFull).
                uint32_t c = c0;
                asm volatile("" : "+v"(c));
                int b = tid;
                for (int k = 0; k < lanesFull; ++k, b += NT, --c) scoreBlockPacked(b, c, false);
                const int extra = nFull - lanesFull * NT;
                if (tid < extra) scoreBlockPacked(tid + lanesFull * NT, c, false);
                if (nFull < nBlk && tid == extra) scoreBlockPacked(nFull, c, true);
            };
            auto select = [&](auto laneTag) {
                constexpr bool kLane = decltype(laneTag)::value;
                const uint32_t pRecipSel = kLane ? 0u : recipOfP();   // slot path: groups by v mod P
                Score accS[8];
                int32_t accV[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) { accS[j] = kNoScore; accV[j] = -1; }
                // int32 scores: the per-group argmax folds packed 64-bit keys
                // across lanes and publishes one per group and wave.
                // Builds with registers to spare run one straight-line fold per
                // group count; the lean 80-VGPR builds one fold with the steps
                // chosen at run time (the ten straight-line copies made them
                // spill to scratch).
                auto foldKeysP = [&](uint64_t (&kk)[8]) {
                    // One straight-line reduction per group count (P a power
                    // of two up to 512): no branches between the steps.
                    auto foldP = [&](auto lgTag) {
                        constexpr int kP = 1 << decltype(lgTag)::value;
                        if constexpr (kP >= 8) {
                            constexpr int Q = kP >> 3;
                            constexpr int held = Q <= 8 ? 1 : Q <= 16 ? 2 : Q <= 32 ? 4 : 8;
                            int base = 0;
                            if constexpr (Q <= 32) {
#pragma unroll
                                for (int j = 0; j < 4; ++j) scatterStepKey<32>(kk[j], kk[j + 4], lane);
                                base += (lane & 32) ? 4 : 0;
                            }
                            if constexpr (Q <= 16) {
#pragma unroll
                                for (int j = 0; j < 2; ++j) scatterStepKey<16>(kk[j], kk[j + 2], lane);
                                base += (lane & 16) ? 2 : 0;
                            }
                            if constexpr (Q <= 8) {
                                scatterStepKey<8>(kk[0], kk[1], lane);
                                base += (lane & 8) ? 1 : 0;
                            }
                            if constexpr (Q <= 4) foldLanesKey<4>(kk[0], lane);
                            if constexpr (Q <= 2) foldLanesKey<2>(kk[0], lane);
                            if constexpr (Q <= 1) foldLanesKey<1>(kk[0], lane);
                            if ((lane & 7) < Q) {
                                const int row = wave * pstride + 8 * (lane & (Q - 1)) + base;
#pragma unroll
                                for (int k = 0; k < held; ++k) part[row + k] = kk[k];
                            }
                        } else {
                            if constexpr (kP <= 4) {
#pragma unroll
                                for (int j = 0; j < 4; ++j) keepMaxKey(kk[j], kk[j + 4]);
                            }
                            if constexpr (kP <= 2) {
#pragma unroll
                                for (int j = 0; j < 2; ++j) keepMaxKey(kk[j], kk[j + 2]);
                            }
                            if constexpr (kP == 1) keepMaxKey(kk[0], kk[1]);
                            int base = 0;
                            if constexpr (kP == 4) {
                                scatterStepKey<32>(kk[0], kk[2], lane);
                                scatterStepKey<32>(kk[1], kk[3], lane);
                                base += (lane & 32) ? 2 : 0;
                                scatterStepKey<16>(kk[0], kk[1], lane);
                                base += (lane & 16) ? 1 : 0;
                            } else if constexpr (kP == 2) {
                                scatterStepKey<32>(kk[0], kk[1], lane);
                                base += (lane & 32) ? 1 : 0;
                            }
                            if constexpr (kP == 1) foldLanesKey<32>(kk[0], lane);
                            if constexpr (kP <= 2) foldLanesKey<16>(kk[0], lane);
                            foldLanesKey<8>(kk[0], lane);
                            foldLanesKey<4>(kk[0], lane);
                            foldLanesKey<2>(kk[0], lane);
                            foldLanesKey<1>(kk[0], lane);
                            if ((lane & (63 >> (kP >> 1))) == 0) part[wave * pstride + base] = kk[0];
                        }
                    };
                    switch (P) {
                        case 1: foldP(std::integral_constant<int, 0>{}); break;
                        case 2: foldP(std::integral_constant<int, 1>{}); break;
                        case 4: foldP(std::integral_constant<int, 2>{}); break;
                        case 8: foldP(std::integral_constant<int, 3>{}); break;
                        case 16: foldP(std::integral_constant<int, 4>{}); break;
                        case 32: foldP(std::integral_constant<int, 5>{}); break;
                        case 64: foldP(std::integral_constant<int, 6>{}); break;
                        case 128: foldP(std::integral_constant<int, 7>{}); break;
                        case 256: foldP(std::integral_constant<int, 8>{}); break;
                        default: foldP(std::integral_constant<int, 9>{}); break;   // 512
                    }
                };
                auto foldKeysRt = [&](uint64_t (&kk)[8]) {
                    if (P >= 8) {
                        const int Q = P >> 3;
                        int held = 8, base = 0;
                        if (Q <= 32) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) scatterStepKey<32>(kk[j], kk[j + 4], lane);
                            held = 4;
                            base += (lane & 32) ? 4 : 0;
                        }
                        if (Q <= 16) {
#pragma unroll
                            for (int j = 0; j < 2; ++j) scatterStepKey<16>(kk[j], kk[j + 2], lane);
                            held = 2;
                            base += (lane & 16) ? 2 : 0;
                        }
                        if (Q <= 8) {
                            scatterStepKey<8>(kk[0], kk[1], lane);
                            held = 1;
                            base += (lane & 8) ? 1 : 0;
                        }
                        if (Q <= 4) foldLanesKey<4>(kk[0], lane);
                        if (Q <= 2) foldLanesKey<2>(kk[0], lane);
                        if (Q <= 1) foldLanesKey<1>(kk[0], lane);
                        if ((lane & 7) < Q) {
                            const int row = wave * pstride + 8 * (lane & (Q - 1)) + base;
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                if (k < held) part[row + k] = kk[k];
                        }
                    } else {
                        if (P <= 4) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) keepMaxKey(kk[j], kk[j + 4]);
                        }
                        if (P <= 2) {
#pragma unroll
                            for (int j = 0; j < 2; ++j) keepMaxKey(kk[j], kk[j + 2]);
                        }
                        if (P == 1) keepMaxKey(kk[0], kk[1]);
                        int base = 0;
                        if (P == 4) {
                            scatterStepKey<32>(kk[0], kk[2], lane);
                            scatterStepKey<32>(kk[1], kk[3], lane);
                            base += (lane & 32) ? 2 : 0;
                            scatterStepKey<16>(kk[0], kk[1], lane);
                            base += (lane & 16) ? 1 : 0;
                        } else if (P == 2) {
                            scatterStepKey<32>(kk[0], kk[1], lane);
                            base += (lane & 32) ? 1 : 0;
                        }
                        if (P == 1) foldLanesKey<32>(kk[0], lane);
                        if (P <= 2) foldLanesKey<16>(kk[0], lane);
                        foldLanesKey<8>(kk[0], lane);
                        foldLanesKey<4>(kk[0], lane);
                        foldLanesKey<2>(kk[0], lane);
                        foldLanesKey<1>(kk[0], lane);
                        if ((lane & (63 >> (P >> 1))) == 0) part[wave * pstride + base] = kk[0];
                    }
                };
                auto foldKeys = [&](uint64_t (&kk)[8]) {
                    if constexpr (kFoldPerP) foldKeysP(kk);
                    else foldKeysRt(kk);
                };
                // int32 scores on the lane path: the lane keys (packed, or
                // (score, block) pairs when the packed bound fails) become
                // 64-bit keys and take the one reduction above.
                uint64_t kk[8];
                bool haveKeys = false;
                if constexpr (kLane && kPackable) {
                    if (packShift >= 0) {
                        // The lane keys fold as they are: the high word is the
                        // key (score, then the earlier block k), biased to
                        // unsigned, and the low word ~(8 tid + j) breaks the
                        // remaining ties toward the smaller lane and slot.
                        // (score, k, tid, j) orders as (score, vertex) does,
                        // since vertex = 8 (tid + k NT) + j; RESOLVE recovers
                        // k from the key's low packShift bits.  No real key is
                        // INT32_MIN (|score| < 2^(31 - shift)): a slot left
                        // there scored nothing, which happens only when some
                        // lane has no full block, and becomes the empty key 0.
                        int32_t acc[8];
                        scorePackedLanes(acc);
                        const bool mayBeEmpty = lanesFull == 0;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            kk[j] = (static_cast<uint64_t>(static_cast<uint32_t>(acc[j]) ^ 0x80000000u) << 32) |
                                    static_cast<uint64_t>(~static_cast<uint32_t>(8 * tid + j));
                            if (mayBeEmpty && acc[j] == INT32_MIN) kk[j] = 0ull;
                        }
                        haveKeys = true;
                    }
                }
                auto scoreBlock = [&](int b, bool tail) {
                    Block8<kBits> blk;
                    blk.template load<kHbm>(wl, b);
                    uint32_t o[4];
                    philoxBlock(phx, pkeys, static_cast<uint32_t>(b), o);
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int v = 8 * b + j;
                        if (tail && v >= n) continue;
                        const uint32_t r = ((j & 1) ? (o[j >> 1] >> 16) : o[j >> 1]);
                        Score score;
                        if constexpr (kWide)
                            score = scoreWide(beta, blk.field(j), lutw[r & 15u], (r >> 4) & 0x7Fu);
                        else
                            score = score32(blk, j, r);
                        if constexpr (kLane) {
                            // slot j keeps the block index; the vertex is 8*b + j
                            if (score > accS[j]) { accS[j] = score; accV[j] = b; }
                        } else {
                            const int g = groupOf(v, pRecipSel);
                            if constexpr (kWide) {
                                if (!idPass)
                                    atomicMax(reinterpret_cast<unsigned long long*>(&part[g]),
                                              static_cast<unsigned long long>(biased64(score)));
                                else if (biased64(score) == part[g])
                                    atomicMax(&partId[g], ~static_cast<uint32_t>(v));
                            } else {
                                atomicMax(reinterpret_cast<unsigned long long*>(&part[g]),
                                          static_cast<unsigned long long>(packKey(score, v)));
                            }
                        }
                    }
                };
                if (!haveKeys) {
                    for (int b = tid; b < nFull; b += NT) scoreBlock(b, false);
                    if (nFull < nBlk && tid == (nFull % NT)) scoreBlock(nFull, true);
                    if constexpr (kLane) {
                        // Block index -> vertex; an empty slot's -1 becomes
                        // j - 8, still negative, so it stays empty.
#pragma unroll
                        for (int j = 0; j < 8; ++j) accV[j] = 8 * accV[j] + j;
                        if constexpr (kKeyFold) {
#pragma unroll
                            for (int j = 0; j < 8; ++j) kk[j] = packKey(accS[j], accV[j]);
                        }
                    }
                }

                if constexpr (kLane) {
                    if (wave >= busyWaves) return;
                    if constexpr (kKeyFold) {
                        foldKeys(kk);
                        return;
                    }
                    if (P >= 8) {
                        // Lanes l and l' share groups iff l = l' (mod Q); slot j
                        // of lane l feeds group 8 (l mod Q) + j.  Reduce-scatter
                        // over the lane distances 32, 16, 8 that are >= Q: each
                        // step halves the slots a lane holds (held slot k is
                        // slot base + k, base from the lane's bits 5, 4, 3),
                        // then plain folds over the distances 4, 2, 1 >= Q.
                        const int Q = P >> 3;
                        int held = 8, base = 0;
                        if (Q <= 32) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) scatterStep<32>(accS[j], accV[j], accS[j + 4], accV[j + 4], lane);
                            held = 4;
                            base += (lane & 32) ? 4 : 0;
                        }
                        if (Q <= 16) {
#pragma unroll
                            for (int j = 0; j < 2; ++j) scatterStep<16>(accS[j], accV[j], accS[j + 2], accV[j + 2], lane);
                            held = 2;
                            base += (lane & 16) ? 2 : 0;
                        }
                        if (Q <= 8) {
                            scatterStep<8>(accS[0], accV[0], accS[1], accV[1], lane);
                            held = 1;
                            base += (lane & 8) ? 1 : 0;
                        }
                        if (Q <= 4) foldLanes<4>(accS[0], accV[0], lane);
                        if (Q <= 2) foldLanes<2>(accS[0], accV[0], lane);
                        if (Q <= 1) foldLanes<1>(accS[0], accV[0], lane);
                        // After the folds the lanes differing only in bits
                        // below 3 and at or above log2 Q hold equal values:
                        // one of them writes.
                        if ((lane & 7) < Q) {
                            const int row = wave * pstride + 8 * (lane & (Q - 1)) + base;
#pragma unroll
                            for (int k = 0; k < 8; ++k)
                                if (k < held) putPartial(row + k, accS[k], accV[k]);
                        }
                    } else {
                        // P in {1, 2, 4}: slot j feeds group j mod P.
                        if (P <= 4) {
#pragma unroll
                            for (int j = 0; j < 4; ++j) keepBetter(accS[j], accV[j], accS[j + 4], accV[j + 4]);
                        }
                        if (P <= 2) {
#pragma unroll
                            for (int j = 0; j < 2; ++j) keepBetter(accS[j], accV[j], accS[j + 2], accV[j + 2]);
                        }
                        if (P == 1) keepBetter(accS[0], accV[0], accS[1], accV[1]);
                        // Every lane now holds slots 0..P-1 (group j); the
                        // same reduce-scatter leaves one group per lane (base),
                        // then folds over the remaining lane bits.
                        int base = 0;
                        if (P == 4) {
                            scatterStep<32>(accS[0], accV[0], accS[2], accV[2], lane);
                            scatterStep<32>(accS[1], accV[1], accS[3], accV[3], lane);
                            base += (lane & 32) ? 2 : 0;
                            scatterStep<16>(accS[0], accV[0], accS[1], accV[1], lane);
                            base += (lane & 16) ? 1 : 0;
                        } else if (P == 2) {
                            scatterStep<32>(accS[0], accV[0], accS[1], accV[1], lane);
                            base += (lane & 32) ? 1 : 0;
                        }
                        if (P == 1) foldLanes<32>(accS[0], accV[0], lane);
                        if (P <= 2) foldLanes<16>(accS[0], accV[0], lane);
                        foldLanes<8>(accS[0], accV[0], lane);
                        foldLanes<4>(accS[0], accV[0], lane);
                        foldLanes<2>(accS[0], accV[0], lane);
                        foldLanes<1>(accS[0], accV[0], lane);
                        // lanes 0, 16, 32, 48 (P = 4), 0 and 32 (P = 2), 0 (P = 1)
                        if ((lane & (63 >> (P >> 1))) == 0) putPartial(wave * pstride + base, accS[0], accV[0]);
                    }
                }
            };
            if (laneReduce) {
                select(std::true_type{});
            } else {
                select(std::false_type{});
                if constexpr (kWide) {
                    __syncthreads();   // every group's max score is in its slot
                    idPass = true;
                    select(std::false_type{});
                }
            }
            __syncthreads();   // B1: all partials / slots written
            const bool packedKeys = kPackable && laneReduce && packShift >= 0;
            const uint32_t packMask = packShift > 0 ? (1u << packShift) - 1u : 0u;

            // ---------------- RESOLVE: winner of each group ----------------
            for (int g = tid; g < P; g += NT) {
                uint64_t key = 0ull;
                uint32_t kid = 0u;   // ~id of the winner (int64 scores)
                // Every entry read here is cleared again, so the region is all
                // zero between sweeps whichever reduction the next phase uses.
                if (laneReduce) {
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        // An idle wave's entries stay zero (empty), so up to
                        // four waves are read without a branch.
                        if constexpr (NW > 4)
                            if (w >= busyWaves) break;
                        const uint64_t k2 = part[w * pstride + g];
                        part[w * pstride + g] = 0ull;
                        if constexpr (kWide) {
                            const uint32_t i2 = partId[w * pstride + g];
                            partId[w * pstride + g] = 0u;
                            if (k2 > key || (k2 == key && i2 > kid)) { key = k2; kid = i2; }
                        } else {
                            key = k2 > key ? k2 : key;
                        }
                    }
                } else {
                    key = part[g];
                    part[g] = 0ull;
                    if constexpr (kWide) {
                        kid = partId[g];
                        partId[g] = 0u;
                    }
                }
                int32_t pick = -1, sp = 0;
                Delta dl = 0;
                if (key != 0ull) {
                    pick = static_cast<int32_t>(~(kWide ? kid : static_cast<uint32_t>(key)));
                    // Lane keys carry 8 tid + j; the block position k is
                    // K - 1 minus the key's low packShift bits.
                    if (packedKeys)
                        pick += (lanesK - 1 - static_cast<int32_t>(static_cast<uint32_t>(key >> 32) & packMask))
                                << (3 + kLgNT);
                    const Word w = ldsWord<kBits, kHbm>(wl, pick);
                    sp = (w & F::kSpin) ? 1 : -1;
                    dl = static_cast<Delta>(w & F::kField) - static_cast<Delta>(F::kBias);
                }
                pks[g] = i32x2{pick, sp};
                pd[g] = dl;
            }
            __syncthreads();   // B2: winners published

            // Beta of the next sweep, and its biased LUT (SELECT reads are done).
            int32_t betaNext = beta;
            if (runStop <= step + 1 && run < nRuns - 1) {
                ++run;
                runStop = static_cast<uint64_t>(uni(nextEnd));
                betaNext = uni(nextVal);
                while (run < nRuns - 1 && runStop <= step + 1) {   // empty runs only
                    ++run;
                    runStop = static_cast<uint64_t>(uni(betaEnd[run]));
                    betaNext = uni(betaVal[run]);
                }
            }
            setLut(betaNext);

            // ---------------- FLIP ----------------
            int64_t dsum = 0;
            {
                // 2^lgL = 2^floor(log2(threads / P)) lanes per winner (at most
                // a wave), from ceil(log2 P) without a division.
                const int ceilLgP = P <= 1 ? 0 : 32 - __clz(P - 1);
                int lgL = kLgNT - ceilLgP < 0 ? 0 : (kLgNT - ceilLgP > 6 ? 6 : kLgNT - ceilLgP);
                // No more lanes than the longest row has edge pairs: they
                // would have nothing to walk, and the winners then fit in
                // fewer waves (the rest leave FLIP at once).
                lgL = lgL < rowLanesLg ? lgL : rowLanesLg;
                const int perPass = NT >> lgL;
                const int passes = (P + perPass - 1) >> (kLgNT - lgL);
                const uint32_t pRecip = recipOfP();
                const int L = 1 << lgL;
                const int sub = tid & (L - 1);
                // A winner's own word takes one atomic per adjacent winner
                // (its coupling) beside the flip itself, instead of a sum
                // over the winner's lanes first, when every partial sum of
                // its update stays inside the field: |delta| + |own| <= 3B.
                const bool splitOwn = 3 * deltaMax <= static_cast<uint64_t>(F::kMaxB);
                auto flipPasses = [&](auto pow2Tag) {
                    constexpr bool kPow2 = decltype(pow2Tag)::value;
                    auto grp = [&](int x) {
                        if constexpr (kPow2) return x & (P - 1);
                        else return modBy(static_cast<uint32_t>(x), static_cast<uint32_t>(P), pRecip);
                    };
                    for (int pass = 0; pass < passes; ++pass) {
                        // Groups rise with the pass and the lane: a wave whose
                        // first group is past P has nothing left to do.
                        if (pass * perPass + ((wave << 6) >> lgL) >= P) break;
                        const int g = pass * perPass + (tid >> lgL);
                        int32_t b = -1, sb = 0;
                        if (g < P) {
                            const i32x2 rec = pks[g];
                            b = rec.x;
                            sb = rec.y;
                        }
                        Delta own = 0;   // sum over adjacent winners x of 2 w s_b s_x (!splitOwn)
                        // Walks row [lo, hi) of winner b, two entries per
                        // lane and step; edgeAt(i) gives entry i as
                        // {neighbour, weight}.  Each entry becomes one atomic
                        // add: to the neighbour x, s_x (-2 w s_b) (x's spin
                        // is stable: only winners flip), or, when x is a
                        // winner too, its coupling 2 w s_b s_x to b's own word
                        // (and to the cut delta when x's group comes first,
                        // as sequential group-order flips see it).  The reads
                        // run ahead of the atomics (a winner's word read here
                        // is not used); a step past the row's end re-reads a
                        // real entry and adds zero.
                        auto walk = [&](uint32_t lo, uint32_t hi, auto edgeAt) {
                            for (uint32_t i0 = lo + sub; i0 < hi; i0 += 2 * L) {
                                const uint32_t i1 = i0 + static_cast<uint32_t>(L);
                                const bool two = i1 < hi;
                                const i32x2 e0 = edgeAt(i0);
                                const i32x2 e1 = edgeAt(two ? i1 : i0);
                                const i32x2 r0 = pks[grp(e0.x)];
                                const i32x2 r1 = pks[grp(e1.x)];
                                const Word wd0 = ldsWord<kBits, kHbm>(wl, e0.x);
                                const Word wd1 = ldsWord<kBits, kHbm>(wl, e1.x);
                                // Spins are +/-1: the products 2 w s_b and
                                // its coupling w2 s_x are sign selections.
                                auto apply = [&](const i32x2& e, const i32x2& r, Word wd, bool live) {
                                    const Delta w = live ? static_cast<Delta>(e.y) : Delta{0};
                                    const Delta w2 = sb > 0 ? w + w : -(w + w);
                                    const bool isW = r.x == e.x;
                                    const Delta t = r.y > 0 ? w2 : -w2;
                                    if (isW && grp(e.x) < g) dsum -= t;
                                    Delta add = (wd & F::kSpin) ? -w2 : w2;
                                    int at = e.x;
                                    if (isW) {
                                        if (splitOwn) at = b; else own += t;
                                        add = splitOwn ? t : Delta{0};
                                    }
                                    ldsAddWord<kBits, kHbm>(wl, at, static_cast<Word>(add));
                                };
                                apply(e0, r0, wd0, true);
                                apply(e1, r1, wd1, two);
                            }
                        };
                        if (b >= 0) {
                            if (rowsInLds) {
                                walk(rowOffL[b], rowOffL[b + 1], [&](uint32_t i) {
                                    const uint32_t c = rowEdgL[i];
                                    return i32x2{static_cast<int32_t>(c & ~kCompactNeg), (c & kCompactNeg) ? -1 : 1};
                                });
                            } else {
                                walk(offsets[b], offsets[b + 1],
                                     [&](uint32_t i) { return static_cast<i32x2>(edges[i]); });
                            }
                        }
                        // Sum over the winner's L lanes (an aligned lane group).
                        if (!splitOwn) own = laneGroupSum(own, L, lane);
                        if (b >= 0 && sub == 0) {
                            dsum += pd[g];
                            // delta(b) -> -delta(b) + own, spin bit toggled: one add.
                            const Word toggle = sb > 0 ? Word{0} - F::kSpin : F::kSpin;
                            ldsAddWord<kBits, kHbm>(
                                wl, b, static_cast<Word>(own) - Word{2} * static_cast<Word>(pd[g]) + toggle);
                        }
                    }
                };
                if ((P & (P - 1)) == 0) flipPasses(std::true_type{});
                else flipPasses(std::false_type{});
            }
            // |cut delta| <= 3 * B * P < 2^31 below 32-bit words.
            if constexpr (kBits < 32)
                dsum = waveSum32(static_cast<int32_t>(dsum));
            else
                dsum = waveSum64(dsum);
            if (lane == 0) dpart[wave] = dsum;
            __syncthreads();   // B3: spins, fields, LUT and partial sums complete
            if constexpr (kHbm) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");

            int64_t total = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) total += dpart[w];
            cur += uni(total);
            if (cur > best) { best = cur; improved = 1; } else { improved = 0; }
            beta = betaNext;
            ++step;
        }

        // The slice's best state: the last sweep's, or the one before it.
        if constexpr (kDeferBest) {
            if (improved && step != stepStart)
                snapshot();
            else if (pending && !improved)
                snapshotBeforeLast();
        } else {
            if (improved) snapshot();
        }
        __syncthreads();
        if constexpr (!kHbm) {
            GlobalOf<u32x4>* gw = globalIn(reinterpret_cast<u32x4*>(it.words));
            const u32x4* lw = reinterpret_cast<const u32x4*>(wl);
            for (int i = tid; i < wq; i += NT) gw[i] = lw[i];
            GlobalOf<u32x4>* gb = globalIn(reinterpret_cast<u32x4*>(it.bestbits));
            const u32x4* lb = reinterpret_cast<const u32x4*>(bl);
            for (int i = tid; i < bq; i += NT) gb[i] = lb[i];
        }
        if (tid == 0) {
            st->cur = cur;
            st->best = best;
            st->step = step;
            st->improved = improved;
        }
        __syncthreads();
    }

    // Last workgroup out resets the work counters, so the next launch (or
    // graph replay) starts from zero without a memset.
    if (tid == 0) {
        __threadfence();
        const uint32_t done = atomicAdd(&sched[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&sched[0], 0u);
            atomicExch(&sched[1], 0u);
        }
    }
}
}  // namespace dev

namespace launch {

// The compiled variants of one word width, score mode and state placement:
// threads per chain, the register-rich builds and the rows-in-LDS builds.
template <int kBits, int kMode, bool kHbm>
const void* sliceKernel(int threads, bool rich, bool rows) {
    if constexpr ((kMode == dev::kScoreTab || kMode == dev::kScore24) && !kHbm) {
        if (rich && threads == 128)
            return rows ? reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 128, kHbm, true, true>)
                        : reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 128, kHbm, true>);
        if (rich && threads == 256)
            return rows ? reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 256, kHbm, true, true>)
                        : reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 256, kHbm, true>);
    }
    if constexpr (!kHbm) {
        if (rows && threads == 512) return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 512, kHbm, false, true>);
        if (rows && threads == 1024) return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 1024, kHbm, false, true>);
    }
    if (rows) return nullptr;
    switch (threads) {
        case 128: return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 128, kHbm>);
        case 256: return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 256, kHbm>);
        case 512: return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 512, kHbm>);
        case 1024: return reinterpret_cast<const void*>(dev::annealSlice<kBits, kMode, 1024, kHbm>);
        default: return nullptr;
    }
}

}  // namespace launch
}  // namespace rfmf
