// anneal_gpu.hpp -- C++ entry point with the signature of the reference's
// gpu3::annealGpu (src/anneal_gpu.cuh:52-88), implemented over the C ABI of
// librfmf.so.  Errors are thrown as std::runtime_error carrying the library
// message (the reference prints and exits instead).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/rfmf.h"
#include "immutable_graph.hpp"

namespace gpu3 {

enum class CorrMode { Walk, Warpwalk };

struct LaunchConfig {
    CorrMode corr = CorrMode::Walk;
    int device = 0;
    int sWidth = 32;
};

inline void annealGpu(const anneallib::AnnealCsr& ig, int64_t maxWeightedDeg,
                      const std::vector<std::vector<int8_t>>& initialSpins,
                      const std::vector<double>& t0s, const std::vector<size_t>& phases,
                      const std::vector<uint64_t>& iters, const std::vector<uint64_t>& solverSeeds,
                      const LaunchConfig& cfg, std::vector<std::vector<int8_t>>& endSpins,
                      std::vector<std::vector<int8_t>>& bestSpins, float* kernelMs,
                      std::vector<int64_t>* kernelBestVals = nullptr,
                      std::vector<int64_t>* kernelFinalVals = nullptr) {
    const size_t C = t0s.size();
    if (C == 0 || solverSeeds.size() != C || initialSpins.size() != C)
        throw std::runtime_error("annealGpu needs non-empty, equal-length t0s (" + std::to_string(t0s.size()) +
                                 "), solverSeeds (" + std::to_string(solverSeeds.size()) +
                                 ") and initialSpins (" + std::to_string(initialSpins.size()) + ")");
    if (phases.size() != iters.size())
        throw std::runtime_error("annealGpu: phases and iters differ in length");
    const size_t n = ig.numVertices();
    std::vector<int8_t> init(C * n), end(C * n), best(C * n);
    for (size_t c = 0; c < C; ++c) {
        if (initialSpins[c].size() != n) throw std::runtime_error("annealGpu: initial spin vector size != n");
        std::copy(initialSpins[c].begin(), initialSpins[c].end(), init.begin() + c * n);
    }
    std::vector<uint64_t> ph(phases.begin(), phases.end());
    rfmf_csr csr{ig.offsets().data(), ig.edgeTo().data(), ig.edgeWeight().data(), static_cast<int64_t>(n),
                 static_cast<int64_t>(ig.edgeTo().size())};
    rfmf_launch_cfg lc{cfg.corr == CorrMode::Warpwalk ? RFMF_CORR_WARPWALK : RFMF_CORR_WALK, cfg.device,
                       cfg.sWidth, 0};
    std::vector<int64_t> bv(C), fv(C);
    const rfmf_status st = rfmf_anneal(&csr, maxWeightedDeg, init.data(), static_cast<int64_t>(C), t0s.data(),
                                       ph.data(), iters.data(), static_cast<int64_t>(ph.size()),
                                       solverSeeds.data(), &lc, end.data(), best.data(), kernelMs, bv.data(),
                                       fv.data());
    if (st != RFMF_OK)
        throw std::runtime_error(std::string("annealGpu: ") + rfmf_status_string(st) + ": " + rfmf_last_error());
    endSpins.assign(C, {});
    bestSpins.assign(C, {});
    for (size_t c = 0; c < C; ++c) {
        endSpins[c].assign(end.begin() + c * n, end.begin() + (c + 1) * n);
        bestSpins[c].assign(best.begin() + c * n, best.begin() + (c + 1) * n);
    }
    if (kernelBestVals) *kernelBestVals = bv;
    if (kernelFinalVals) *kernelFinalVals = fv;
}

}  // namespace gpu3
