// pp_zsolve_pcr.hip -- launches of the PCR -> Thomas z solve (k_zsolve).  One instantiation
// per rows-per-lane J = ceil(nz / 16), so the padding below the last row stays under 16
// rows: the kernel keeps explicit coefficients only for rows j in {0, J-2, J-1}, which
// covers the bottom 16 real rows exactly when the padding is shorter than one lane row.
#include "pp_zsolve.h"

namespace pp {

template <typename R, int J>
static hipError_t launch_pcr(const ZArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.SX + 15) / 16), (unsigned)a.SY);
    const bool full = (a.nz == 16 * J);
    const bool off32 = (double)a.SX * a.SY * a.nz * sizeof(R) < 4294967296.0;
#define PP_ZL(F, O) hipLaunchKernelGGL((k_zsolve<R, J, F, O>), grid, dim3(256), 0, s, a)
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
