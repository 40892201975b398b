// pp_line_pp.hip -- instantiates the PP line transforms (all compiled sizes, x and y axes,
// forward and inverse, FP64 and FP32 forms, and the mixed form's fused residual + x forward).  One translation unit per transform kind so they build in parallel.
#include "pp_line.h"
#include "pp_resx.h"

namespace pp {
template hipError_t launch_line_kind<KIND_PP>(int, int, int, int, const LineArgs &, hipStream_t);
template hipError_t launch_resx_kind<KIND_PP>(int, int, const LineArgs &, const ResArgs &, const ResArgs *,
                                                  hipStream_t);
}  // namespace pp
