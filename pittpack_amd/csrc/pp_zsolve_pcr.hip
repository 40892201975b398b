// pp_zsolve_pcr.hip -- launches of the PCR -> Thomas z solve (k_zsolve).  One instantiation
// per rows-per-lane J = ceil(nz / 16), so the padding below the last row stays under 16
// rows: the kernel keeps explicit coefficients only for rows j in {0, J-2, J-1}, which
// covers the bottom 16 real rows exactly when the padding is shorter than one lane row.
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
static hipError_t launch_pcr_c(const ZArgs &a, hipStream_t s);

template <typename R, int J>
static hipError_t launch_pcr(const ZArgs &a, hipStream_t s)
{
    if (zcpb() == 8) return launch_pcr_c<R, J, 8>(a, s);
    return launch_pcr_c<R, J, 16>(a, s);
}

template <typename R, int J, int CPB>
static hipError_t launch_pcr_c(const ZArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.SX + CPB - 1) / CPB), (unsigned)a.SY);
    const bool full = (a.nz == 16 * J);
    const bool off32 = (double)a.SX * a.SY * a.nz * sizeof(R) < 4294967296.0;
    constexpr size_t lds = ztile_bytes_cyclic<R, J, CPB>();
#define PP_ZL(F, O)                                                                                            \
    do {                                                                                                       \
        auto kern = k_zsolve<R, J, F, O, CPB>;                                                                            \
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
    if (full) { if (off32) PP_ZL(true, true); else PP_ZL(true, false); }
    else { if (off32) PP_ZL(false, true); else PP_ZL(false, false); }
#undef PP_ZL
    return hipGetLastError();
}

bool zsolve_supported_f64(int nz) { return nz >= 4 && nz <= 16 * 32; }

hipError_t launch_zsolve_pcr_f64(int nz, const ZArgs &a, hipStream_t s)
{
    const int J = (nz + 15) / 16;
    switch (J) {
#define PPJ(v) \
    case v: return launch_pcr<double, v>(a, s);
        PPJ(1) PPJ(2) PPJ(3) PPJ(4) PPJ(5) PPJ(6) PPJ(7) PPJ(8) PPJ(9) PPJ(10) PPJ(11) PPJ(12)
        PPJ(13) PPJ(14) PPJ(15) PPJ(16) PPJ(17) PPJ(18) PPJ(19) PPJ(20) PPJ(21) PPJ(22) PPJ(23)
        PPJ(24) PPJ(25) PPJ(26) PPJ(27) PPJ(28) PPJ(29) PPJ(30) PPJ(31) PPJ(32)
#undef PPJ
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pp
