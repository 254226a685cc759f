// slice_w64_score64.hip -- annealSlice variants for 64-bit words and int64 scores
// (state in LDS and in HBM).
#include "anneal_slice.hpp"

namespace rfmf {
namespace launch {

template const void* sliceKernel<64, dev::kScore64, false>(int, bool, bool);
template const void* sliceKernel<64, dev::kScore64, true>(int, bool, bool);

}  // namespace launch
}  // namespace rfmf
