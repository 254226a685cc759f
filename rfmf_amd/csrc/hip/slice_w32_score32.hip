// slice_w32_score32.hip -- annealSlice variants for 32-bit words and int32 scores
// (state in LDS and in HBM).
#include "anneal_slice.hpp"

namespace rfmf {
namespace launch {

template const void* sliceKernel<32, dev::kScore32, false>(int, bool, bool);
template const void* sliceKernel<32, dev::kScore32, true>(int, bool, bool);

}  // namespace launch
}  // namespace rfmf
