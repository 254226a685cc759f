// slice_w16_score32.hip -- annealSlice variants for 16-bit words and int32 scores
// (state in LDS and in HBM).
#include "anneal_slice.hpp"

namespace rfmf {
namespace launch {

template const void* sliceKernel<16, dev::kScore32, false>(int, bool, bool);
template const void* sliceKernel<16, dev::kScore32, true>(int, bool, bool);

}  // namespace launch
}  // namespace rfmf
