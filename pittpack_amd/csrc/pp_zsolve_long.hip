// pp_zsolve_long.hip -- PCR -> Thomas z solve (k_zsolve) for FP64 columns of 513 to 1024 rows,
// J = ceil(nz / 16) = 33..64 rows per lane.  Blocks take 8 columns (a 16 x 1024 FP64 tile
// would need 139 KB of LDS and leave one block per CU; 8 columns fit two), numbered by XCD so
// the two blocks sharing each 128-byte row sit on one L2; the Thomas factors live in the
// lanes' tile slots.  The field is read in place or from the y transform's z tiles (8 or 16
// columns wide: a block takes all or half of a tile).
#include "pp_launch.h"
#include "pp_zsolve.h"

namespace pp {

template <int J, bool FULL>
static hipError_t launch_long(const ZArgs &a, hipStream_t s)
{
    constexpr int CPB = 8;
    constexpr size_t lds = ztile_bytes_cyclic<double, J, CPB>();
    auto kern = k_zsolve<double, J, FULL, true, CPB>;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    const dim3 grid((unsigned)((a.SX + CPB - 1) / CPB), (unsigned)a.SY);
    hipLaunchKernelGGL(kern, grid, dim3(16 * CPB), lds, s, a);
    return hipGetLastError();
}

bool zsolve_long_supported_f64(int nz) { return nz > 16 * 32 && nz <= 16 * 64; }

hipError_t launch_zsolve_long_f64(int nz, const ZArgs &a, hipStream_t s)
{
    // every column is written back (no mixed-form skipping); in place, rows are addressed with
    // 32-bit lane offsets (16 rows of the field below 4 GiB); z tiles are 8 or 16 columns wide
    if (!zsolve_long_supported_f64(nz) || a.nz != nz || a.items || a.tau != 0.0) return hipErrorInvalidValue;
    if (a.tiled && (a.tile_w <= 0 || a.tile_w % 8 != 0)) return hipErrorInvalidValue;
    if (!a.tiled && 16.0 * a.SX * a.SY * sizeof(double) >= 4294967296.0) return hipErrorInvalidValue;
    if (nz == 16 * 64) return launch_long<64, true>(a, s);
    switch ((nz + 15) / 16) {
#define PPL(v) \
    case v: return launch_long<v, false>(a, s);
        PPL(33) PPL(34) PPL(35) PPL(36) PPL(37) PPL(38) PPL(39) PPL(40) PPL(41) PPL(42) PPL(43) PPL(44) PPL(45)
        PPL(46) PPL(47) PPL(48) PPL(49) PPL(50) PPL(51) PPL(52) PPL(53) PPL(54) PPL(55) PPL(56) PPL(57) PPL(58)
        PPL(59) PPL(60) PPL(61) PPL(62) PPL(63) PPL(64)
#undef PPL
    default: return hipErrorInvalidValue;
    }
}

}  // namespace pp
