// slice_w8_table.hip -- annealSlice variants for 8-bit words and the static score table
// (state in LDS).
#include "anneal_slice.hpp"

namespace rfmf {
namespace launch {

template const void* sliceKernel<8, dev::kScoreTab, false>(int, bool, bool);

}  // namespace launch
}  // namespace rfmf
