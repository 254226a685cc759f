// slice_w8_score24.hip -- annealSlice variants for 8-bit words and 24-bit-multiply int32 scores
// (state in LDS and in HBM).
#include "anneal_slice.hpp"

namespace rfmf {
namespace launch {

template const void* sliceKernel<8, dev::kScore24, false>(int, bool, bool);
template const void* sliceKernel<8, dev::kScore24, true>(int, bool, bool);

}  // namespace launch
}  // namespace rfmf
