// pp_line_bs_dd.hip -- instantiates the arbitrary-length DD line transforms (pp_line_bs.h): every
// convolution length, x and y axes, forward and inverse, FP64 and FP32 forms.
#include "pp_line_bs.h"

namespace pp {
template hipError_t launch_line_bs_kind<KIND_DD>(int, int, int, const LineArgs &, hipStream_t);
}  // namespace pp
