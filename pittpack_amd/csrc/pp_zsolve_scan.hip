// pp_zsolve_scan.hip -- launches of the partitioned Thomas z solve (k_zsolve_scan).  Any
// padding below the last row is exact for this solver, so J is rounded up to a compiled
// value.
#include <cstdlib>

#include "pp_zsolve.h"

namespace pp {

// columns per block: 8 (64-byte rows, 4 blocks per CU) or 16 (128-byte rows);
// PP_ZCPB in the environment overrides for measurement
static int zcpb()
{
    static int v = [] {
        const char *e = getenv("PP_ZCPB");
        int x = e ? atoi(e) : 16;
        return (x == 8) ? 8 : 16;
    }();
    return v;
}

template <typename R, int J, int CPB>
static hipError_t launch_scan_c(const ZArgs &a, hipStream_t s);

template <typename R, int J>
static hipError_t launch_scan(const ZArgs &a, hipStream_t s)
{
    if (zcpb() == 8) return launch_scan_c<R, J, 8>(a, s);
    return launch_scan_c<R, J, 16>(a, s);
}

template <typename R, int J, int CPB>
static hipError_t launch_scan_c(const ZArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.SX + CPB - 1) / CPB), (unsigned)a.SY);
    const bool full = (a.nz == 16 * J);
    if (16.0 * a.SX * a.SY * sizeof(R) >= 4294967296.0) return hipErrorInvalidValue;
    constexpr size_t lds = ztile_bytes_blocked<R, J, CPB>();
#define PP_ZL(F, O)                                                                                            \
    do {                                                                                                       \
        auto kern = k_zsolve_scan<R, J, F, O, CPB>;                                                                            \
        if (lds > 64 * 1024) {                                                                                 \
            static bool attr = false;                                                                          \
            if (!attr) {                                                                                       \
                hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); \
                if (e != hipSuccess) return e;                                                                 \
                attr = true;                                                                                   \
            }                                                                                                  \
        }                                                                                                      \
        hipLaunchKernelGGL(kern, grid, dim3(16 * CPB), lds, s, a);                                                  \
    } while (0)
    if (full) PP_ZL(true, true);
    else PP_ZL(false, true);
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
