// anneal_kernels.hip -- launchers of the annealing kernels: the slice
// kernel's variant table (variants compiled in the slice_w*.hip units), the
// initial-state kernel and the plain copy kernel used to calibrate HBM
// counters.
#include <hip/hip_runtime.h>

#include "anneal_device.hpp"
#include "anneal_launch.hpp"
#include "init_words.hpp"

namespace rfmf {
namespace dev {

// Packs initial spins into state words and best bits (initWordGroup), one
// thread per 32 vertices, chain blockIdx.y of the item array.
template <int kBits>
__global__ __launch_bounds__(256) void initWords(const InitItem* __restrict__ items, int nItems) {
    const InitItem& it = items[blockIdx.y];
    const int nGroups = (it.n + 31) >> 5;
    for (int grp = blockIdx.x * blockDim.x + threadIdx.x; grp < nGroups; grp += gridDim.x * blockDim.x)
        (void)initWordGroup<kBits>(it.offsets, it.edges, it.spins, it.words, it.bestbits, it.n, grp);
}

// Plain device copy (16 B per lane, grid-stride): the known-size reference
// kernel that HBM traffic counters are calibrated against.
__global__ __launch_bounds__(256) void streamCopy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                  size_t n16) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n16;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

}  // namespace dev

namespace launch {

using dev::ChainItem;
using dev::InitItem;
using dev::LdsPlan;

hipError_t streamCopy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(dev::streamCopy, dim3(4096), dim3(256), 0, s, static_cast<const uint4*>(src),
                       static_cast<uint4*>(dst), n16);
    return hipGetLastError();
}

template <int kBits, bool kHbm>
static const void* kernelForMode(int mode, int threads, bool rich, bool rows) {
    switch (mode) {
        case dev::kScore32:   // int32 scores up to 32-bit words
            if constexpr (kBits <= 32) return sliceKernel<kBits, dev::kScore32, kHbm>(threads, false, rows);
            return nullptr;
        case dev::kScore24:   // 24-bit fields only (8/16-bit words)
            if constexpr (kBits <= 16) return sliceKernel<kBits, dev::kScore24, kHbm>(threads, rich, rows);
            return nullptr;
        case dev::kScore64: return sliceKernel<kBits, dev::kScore64, kHbm>(threads, false, rows);
        case dev::kScoreTab:   // 8-bit words with LDS-resident state only
            if constexpr (kBits == 8 && !kHbm) return sliceKernel<kBits, dev::kScoreTab, kHbm>(threads, rich, rows);
            return nullptr;
        default: return nullptr;
    }
}

static const void* kernelFor(int bits, int mode, int threads, bool hbm, bool rich, bool rows) {
    switch (bits) {
        case 8:  return hbm ? kernelForMode<8, true>(mode, threads, rich, rows) : kernelForMode<8, false>(mode, threads, rich, rows);
        case 16: return hbm ? kernelForMode<16, true>(mode, threads, rich, rows) : kernelForMode<16, false>(mode, threads, rich, rows);
        case 32: return hbm ? kernelForMode<32, true>(mode, threads, rich, rows) : kernelForMode<32, false>(mode, threads, rich, rows);
        case 64: return hbm ? kernelForMode<64, true>(mode, threads, rich, rows) : kernelForMode<64, false>(mode, threads, rich, rows);
        default: return nullptr;
    }
}

hipError_t anneal(int bits, int mode, int threads, bool hbm, bool rich, bool rows, const ChainItem* items,
                  int nItems, uint32_t* sched, const LdsPlan& plan, int grid, uint64_t sliceSweeps, hipStream_t s) {
    if (rows != (plan.rowsOn != 0)) return hipErrorInvalidValue;
    const void* kern = kernelFor(bits, mode, threads, hbm, rich, rows);
    if (!kern) return hipErrorInvalidValue;
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(plan.total));
    if (e != hipSuccess) return e;
    LdsPlan p = plan;
    void* args[] = {(void*)&items, (void*)&nItems, (void*)&sched, (void*)&p, (void*)&sliceSweeps};
    return hipLaunchKernel(kern, dim3(grid), dim3(threads), args, plan.total, s);
}

int maxActiveBlocks(int bits, int mode, int threads, bool hbm, bool rich, bool rows, size_t lds) {
    const void* kern = kernelFor(bits, mode, threads, hbm, rich, rows);
    if (!kern) return 0;
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) !=
        hipSuccess)
        return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, threads, lds) != hipSuccess) return 0;
    return blocks;
}

hipError_t initWords(int bits, const InitItem* items, int nItems, int maxN, hipStream_t s) {
    if (nItems <= 0) return hipSuccess;
    const int groups = (maxN + 31) / 32;
    const int bx = (groups + 255) / 256;
    dim3 grid(bx > 0 ? bx : 1, nItems);
    switch (bits) {
        case 8:  hipLaunchKernelGGL(dev::initWords<8>, grid, dim3(256), 0, s, items, nItems); break;
        case 16: hipLaunchKernelGGL(dev::initWords<16>, grid, dim3(256), 0, s, items, nItems); break;
        case 32: hipLaunchKernelGGL(dev::initWords<32>, grid, dim3(256), 0, s, items, nItems); break;
        case 64: hipLaunchKernelGGL(dev::initWords<64>, grid, dim3(256), 0, s, items, nItems); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace launch
}  // namespace rfmf
