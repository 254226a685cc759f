#include "max_cut.hpp"

namespace anneallib {
namespace solver {

int64_t cutValueRaw(const uint32_t* offsets, const int32_t* to, const int32_t* w, int64_t n,
                    const int8_t* spin) {
    int64_t cut = 0;
    for (int64_t v = 0; v < n; ++v) {
        const int8_t sv = spin[v];
        for (uint32_t i = offsets[v]; i < offsets[v + 1]; ++i) {
            const int64_t x = to[i];
            cut += (x > v && spin[x] != sv) ? static_cast<int64_t>(w[i]) : 0;
        }
    }
    return cut;
}

int64_t cutValueCsr(const AnnealCsr& ig, const std::vector<int8_t>& spin) {
    return cutValueRaw(ig.offsets().data(), ig.edgeTo().data(), ig.edgeWeight().data(),
                       static_cast<int64_t>(ig.numVertices()), spin.data());
}

}  // namespace solver
}  // namespace anneallib
