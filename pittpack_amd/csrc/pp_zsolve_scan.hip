// pp_zsolve_scan.hip -- launches of the partitioned Thomas z solve (k_zsolve_scan).  Any
// padding below the last row is exact for this solver, so J is rounded up to a compiled
// value.
#include "pp_zsolve.h"

namespace pp {

template <typename R, int J>
static hipError_t launch_scan(const ZArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.SX + 15) / 16), (unsigned)a.SY);
    const bool full = (a.nz == 16 * J);
    const bool off32 = (double)a.SX * a.SY * a.nz * sizeof(R) < 4294967296.0;
#define PP_ZL(F, O) hipLaunchKernelGGL((k_zsolve_scan<R, J, F, O>), grid, dim3(256), 0, s, a)
    if (full) { if (off32) PP_ZL(true, true); else PP_ZL(true, false); }
    else { if (off32) PP_ZL(false, true); else PP_ZL(false, false); }
#undef PP_ZL
    return hipGetLastError();
}

hipError_t launch_zsolve_scan_f64(int nz, const ZArgs &a, hipStream_t s)
{
    const int need = (nz + 15) / 16;
    static constexpr int kJ[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32};
    int J = 0;
    for (int j : kJ)
        if (j >= need) {
            J = j;
            break;
        }
    switch (J) {
#define PPJ(v) \
    case v: return launch_scan<double, v>(a, s);
        PPJ(1) PPJ(2) PPJ(3) PPJ(4) PPJ(5) PPJ(6) PPJ(7) PPJ(8) PPJ(10) PPJ(12) PPJ(14) PPJ(16)
        PPJ(20) PPJ(24) PPJ(28) PPJ(32)
#undef PPJ
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pp
