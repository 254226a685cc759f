// tuning.hpp -- gpu3::tuning: the launch-geometry defaults under the names of
// the reference's compile-time table (src/tuning.hpp:8-18, read by
// src/anneal_kernels.cu:641-661 and main.cu:44-45).
//
// Here they are where the geometry choice starts, not where it ends: the
// autotuner (rfmf_tuning.autotune, always on in the one-shot entry points)
// times the candidate tiles on the live chains and keeps the fastest, and
// rfmf_tuning.threads_per_chain overrides both.  Every value only changes
// timing, never a chain's result.  The values are MI355X's (256 CUs, 64-wide
// waves, a chain's state in LDS), not the V100 record the reference's table
// holds.
#pragma once

#include <cstdint>

namespace gpu3 {
namespace tuning {

// Mean degree (2m / n) from which a caller labels the run with the
// warp-walk corr mode, as main.cu:44-45 does.  The HIP kernel has one FLIP
// walk for every degree, so the mode names the run and changes nothing else.
inline constexpr double corrWarpwalkMinDeg = 8.0;

// Co-resident regime: at least coResidentMinChainsPerCu chains per CU.  The
// 2-wave tile for graphs of at most blocked8CoResBigMinN vertices (and for
// chains whose state lives in HBM), the 4-wave tile above it.
inline constexpr double  coResidentMinChainsPerCu  = 2.0;
inline constexpr int     tpbCoResidentBlocked8     = 128;
inline constexpr int     tpbCoResidentBlocked8BigN = 256;
inline constexpr int64_t blocked8CoResBigMinN      = 2048;

// Latency regime: about one chain per CU or fewer.  8-wave tiles for chains
// in LDS at every n (16 waves add more per-wave work than they hide), for
// either corr label; 16-wave tiles for chains whose state lives in HBM,
// whose loads want more waves in flight.
inline constexpr int     tpbWarpwalk      = 512;
inline constexpr int64_t tpbWalkSmallMaxN = 1024;
inline constexpr int     tpbWalkSmall     = 512;
inline constexpr int     tpbWalkLarge     = 512;
inline constexpr int     tpbStateInHbm    = 1024;

}  // namespace tuning
}  // namespace gpu3
