// max_cut.hpp -- independent host recompute of a cut, API of the reference's
// anneallib::solver::cutValueCsr (src/max_cut.hpp:16): the device tracks the
// cut incrementally, and every reported value is recomputed here from the
// returned spins.
#pragma once

#include <cstdint>
#include <vector>

#include "immutable_graph.hpp"

namespace anneallib {
namespace solver {

// Sum of w over undirected edges (v, to), to > v, whose endpoints differ.
int64_t cutValueCsr(const AnnealCsr& ig, const std::vector<int8_t>& spin);
int64_t cutValueRaw(const uint32_t* offsets, const int32_t* to, const int32_t* w, int64_t n,
                    const int8_t* spin);

}  // namespace solver
}  // namespace anneallib
