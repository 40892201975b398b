// pp_line_dn.hip -- instantiates the DN line transforms (all compiled sizes, x and y axes,
// forward and inverse).  One translation unit per transform kind so they build in parallel.
#include "pp_line.h"

namespace pp {
template hipError_t launch_line_kind<double, KIND_DN>(int, int, int, const LineArgs &, hipStream_t);
}  // namespace pp
