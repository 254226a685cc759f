// rng.hpp -- xoshiro256++ seeded by splitmix64 (Vigna), the generator of the
// initial spins (API of the reference's anneallib::Rng, src/rng.hpp:11-58).
// The annealing itself draws from the counter-based Philox stream on the
// device, never from this.
#pragma once

#include <cstdint>

namespace anneallib {

class Rng {
public:
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (uint64_t& w : s_) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            w = x ^ (x >> 31);
        }
    }

    // n (1..32) bits from a 64-bit reservoir, least significant first; a
    // reservoir holding fewer than n bits is discarded and refilled.
    uint32_t bits(int n) {
        if (left_ < n) {
            reservoir_ = next();
            left_ = 64;
        }
        const uint32_t r = static_cast<uint32_t>(reservoir_ & ((uint64_t{1} << n) - 1));
        reservoir_ >>= n;
        left_ -= n;
        return r;
    }

    uint64_t next() {
        const uint64_t out = rotl(s_[0] + s_[3], 23) + s_[0];
        const uint64_t t = s_[1] << 17;
        s_[2] ^= s_[0];
        s_[3] ^= s_[1];
        s_[1] ^= s_[2];
        s_[0] ^= s_[3];
        s_[2] ^= t;
        s_[3] = rotl(s_[3], 45);
        return out;
    }

private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t s_[4];
    uint64_t reservoir_ = 0;
    int left_ = 0;
};

}  // namespace anneallib
