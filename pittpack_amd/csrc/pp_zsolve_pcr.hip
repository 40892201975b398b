// pp_zsolve_pcr.hip -- launches of the PCR -> Thomas z solve (k_zsolve).  One instantiation
// per rows-per-lane J = ceil(nz / 16), so the padding below the last row stays under 16
// rows: the kernel keeps explicit coefficients only for rows j in {0, J-2, J-1}, which
// covers the bottom 16 real rows exactly when the padding is shorter than one lane row.

#include "pp_launch.h"
#include "pp_zsolve.h"

namespace pp {

// columns per block: 128-byte rows of the column tile (16 FP64 or 32 FP32 columns), so
// every global row access covers whole cache lines and no line is shared by two blocks
// (which run on different XCDs, each with its own L2)
template <typename R, int J, int CPB>
static hipError_t launch_pcr_c(const ZArgs &a, hipStream_t s);

template <typename R, int J>
static hipError_t launch_pcr(const ZArgs &a, hipStream_t s)
{
    return launch_pcr_c<R, J, zsolve_cpb<R>()>(a, s);
}

template <typename R, int J, int CPB>
static hipError_t launch_pcr_c(const ZArgs &a, hipStream_t s)
{
    if (a.tiled && a.tile_w % CPB != 0) return hipErrorInvalidValue;
    dim3 grid((unsigned)((a.SX + CPB - 1) / CPB), (unsigned)a.SY);
    bool full = (a.nz == 16 * J);
    if (a.items) {
        // batched: one block per item, column lengths vary inside the launch (every solve's
        // 16 * SX * SY * sizeof(R) is checked below 4 GiB by the caller)
        grid = dim3((unsigned)a.nitems);
        full = false;
    } else if (16.0 * a.SX * a.SY * sizeof(R) >= 4294967296.0) {
        // staged tiles address rows from a uniform base; lane offsets span 16 rows
        return hipErrorInvalidValue;
    }
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
    if (full) PP_ZL(true, true);
    else PP_ZL(false, true);
#undef PP_ZL
    return hipGetLastError();
}

bool zsolve_supported_f64(int nz) { return nz >= 4 && nz <= 16 * 32; }

#define PPJ32(R)                                                                                        \
    PPJ(R, 1) PPJ(R, 2) PPJ(R, 3) PPJ(R, 4) PPJ(R, 5) PPJ(R, 6) PPJ(R, 7) PPJ(R, 8) PPJ(R, 9) PPJ(R, 10) \
    PPJ(R, 11) PPJ(R, 12) PPJ(R, 13) PPJ(R, 14) PPJ(R, 15) PPJ(R, 16) PPJ(R, 17) PPJ(R, 18)            \
    PPJ(R, 19) PPJ(R, 20) PPJ(R, 21) PPJ(R, 22) PPJ(R, 23) PPJ(R, 24) PPJ(R, 25) PPJ(R, 26)            \
    PPJ(R, 27) PPJ(R, 28) PPJ(R, 29) PPJ(R, 30) PPJ(R, 31) PPJ(R, 32)
#define PPJ(R, v) \
    case v: return launch_pcr<R, v>(a, s);

hipError_t launch_zsolve_pcr_f64(int nz, const ZArgs &a, hipStream_t s)
{
    switch ((nz + 15) / 16) {
        PPJ32(double)
    default: return launch_zsolve_long_f64(nz, a, s);
    }
}

bool zsolve_supported_f32(int nz) { return nz >= 4 && nz <= 16 * 64; }

// FP64 re-solve of the near-singular columns of FP32 data: one block per listed 16-column
// tile (a.items / a.solves / a.nitems), FP64 PCR -> Thomas, only those columns written
template <int J>
static hipError_t launch_fix(const ZArgs &a, hipStream_t s)
{
    constexpr int CPB = 16;
    constexpr size_t lds = ztile_bytes_cyclic<double, J, CPB>();
    auto kern = k_zsolve<double, J, false, true, CPB, float, true>;
    if (lds > 64 * 1024) {
        static bool attr = false;
        if (!attr) {
            hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr = true;
        }
    }
    if (a.nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(kern, dim3((unsigned)a.nitems), dim3(16 * CPB), lds, s, a);
    return hipGetLastError();
}

bool zfix_tiles_supported(int nz) { return nz >= 4 && nz <= 16 * 32; }

hipError_t launch_zfix_tiles_f32(int nz_max, const ZArgs &a, hipStream_t s)
{
    switch ((nz_max + 15) / 16) {
#define PPF(v) \
    case v: return launch_fix<v>(a, s);
        PPF(1) PPF(2) PPF(3) PPF(4) PPF(5) PPF(6) PPF(7) PPF(8) PPF(9) PPF(10) PPF(11) PPF(12) PPF(13) PPF(14) PPF(15)
        PPF(16) PPF(17) PPF(18) PPF(19) PPF(20) PPF(21) PPF(22) PPF(23) PPF(24) PPF(25) PPF(26) PPF(27) PPF(28)
        PPF(29) PPF(30) PPF(31) PPF(32)
#undef PPF
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zsolve_pcr_f32(int nz, const ZArgs &a, hipStream_t s)
{
    switch ((nz + 15) / 16) {
        PPJ32(float)
        PPJ(float, 33) PPJ(float, 34) PPJ(float, 35) PPJ(float, 36) PPJ(float, 37) PPJ(float, 38)
        PPJ(float, 39) PPJ(float, 40) PPJ(float, 41) PPJ(float, 42) PPJ(float, 43) PPJ(float, 44)
        PPJ(float, 45) PPJ(float, 46) PPJ(float, 47) PPJ(float, 48) PPJ(float, 49) PPJ(float, 50)
        PPJ(float, 51) PPJ(float, 52) PPJ(float, 53) PPJ(float, 54) PPJ(float, 55) PPJ(float, 56)
        PPJ(float, 57) PPJ(float, 58) PPJ(float, 59) PPJ(float, 60) PPJ(float, 61) PPJ(float, 62)
        PPJ(float, 63) PPJ(float, 64)
    default: return hipErrorInvalidValue;
    }
}
#undef PPJ
#undef PPJ32

}  // namespace pp

namespace pp {

// The reference's Thomas sweep (thomasLowMem, src/triDiag.cpp:1693-1822, not the pinned
// branch) on one listed FP64 column per block.  Its factors depend only on the column's
// eigenvalue and end rows, so the host computes them in the reference's order (bet_0 = b0,
// tmp_j = 1 / bet_{j-1}, bet_j = eig - tmp_j, the last row bn - tmp) and hands over
// tab = {1 / bet_j, tmp_j} per column.  The two sweeps are then linear recurrences in the
// column's values -- forward x_j = (x_j - x_{j-1}) / bet_j, backward x_j -= tmp_{j+1} x_{j+1}
// -- which one wave runs as a scan: each lane composes the affine maps of its contiguous chunk
// of rows (C = ceil(nz / 64) <= 16), a Kogge-Stone scan across the 64 lanes gives each lane the
// value entering its chunk, and the lane sweeps its rows (about 40 dependent steps per sweep
// instead of nz).  Against the serial sweep this reassociation moves a column by a few 1e-14
// of its largest value; the PCR steps it replaces lose ~3e-12 on such columns.  The PCR
// kernel left these columns untouched (|eig + 2| below its tau).
constexpr int ZTH_CMAX = 16;  // rows per lane: nz <= 64 * 16

__global__ void __launch_bounds__(64) k_zthomas64(ZArgs a0, const ZThomasItem *items, const double *tab)
{
    constexpr int W = 64;
    const ZThomasItem it = items[blockIdx.x];
    ZArgs a = a0;
    if (it.solve >= 0) {
        // batched: the member's field (its z tiles), sizes
        const ZSolve sv = a0.solves[it.solve];
        a.data = reinterpret_cast<double *>(a0.data) + sv.offset;
        a.SX = sv.SX;
        a.SY = sv.SY;
        a.nz = sv.nz;
    }
    const int n = a.nz;
    // element j of the column: in place (stride SX*SY) or in its z tile (stride W)
    long zs = (long)a.SX * a.SY;
    long off = (long)it.sy * a.SX + it.sx;
    if (a.tiled) {
        const int Wt = a.tile_w;
        const int rows = a.tile_rows ? a.tile_rows : (n + 15) / 16 * 16;
        off = (((long)(it.sx / Wt) * a.SY + it.sy) * rows) * Wt + it.sx % Wt;
        zs = Wt;
    }
    double *data = reinterpret_cast<double *>(a.data);
    const double *ib = tab + it.toff, *tp = ib + n;
    const int L = threadIdx.x, C = (n + W - 1) / W, j0 = L * C;
    double x[ZTH_CMAX], g[ZTH_CMAX], q[ZTH_CMAX];
#pragma unroll
    for (int k = 0; k < ZTH_CMAX; k++) {
        const int j = j0 + k;
        const bool in = k < C && j < n;
        x[k] = in ? data[off + (long)j * zs] : 0.0;
        g[k] = in ? ib[j] : 0.0;
        q[k] = (in && j + 1 < n) ? tp[j + 1] : 0.0;
    }
    // forward: x_j = -g_j x_{j-1} + x_j g_j, composed over the chunk, scanned in lane order
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int k = 0; k < ZTH_CMAX; k++) {
        if (k < C && j0 + k < n) {
            const double m = -g[k];
            B = m * B + x[k] * g[k];
            A = m * A;
        }
    }
    for (int s = 1; s < W; s <<= 1) {
        const double Au = __shfl_up(A, s, W), Bu = __shfl_up(B, s, W);
        if (L >= s) {
            B = A * Bu + B;
            A = A * Au;
        }
    }
    double v = __shfl_up(B, 1, W);
    if (L == 0) v = 0.0;
#pragma unroll
    for (int k = 0; k < ZTH_CMAX; k++) {
        if (k < C && j0 + k < n) {
            v = (x[k] - v) * g[k];
            x[k] = v;
        }
    }
    // backward: x_j = -tmp_{j+1} x_{j+1} + x_j, composed over the chunk from its last row,
    // scanned in reverse lane order
    A = 1.0;
    B = 0.0;
#pragma unroll
    for (int k = ZTH_CMAX - 1; k >= 0; k--) {
        if (k < C && j0 + k < n) {
            const double m = -q[k];
            B = m * B + x[k];
            A = m * A;
        }
    }
    for (int s = 1; s < W; s <<= 1) {
        const double Ad = __shfl_down(A, s, W), Bd = __shfl_down(B, s, W);
        if (L + s < W) {
            B = A * Bd + B;
            A = A * Ad;
        }
    }
    v = __shfl_down(B, 1, W);
    if (L == W - 1) v = 0.0;
#pragma unroll
    for (int k = ZTH_CMAX - 1; k >= 0; k--) {
        if (k < C && j0 + k < n) {
            v = x[k] - q[k] * v;
            x[k] = v;
        }
    }
#pragma unroll
    for (int k = 0; k < ZTH_CMAX; k++) {
        const int j = j0 + k;
        if (k < C && j < n) data[off + (long)j * zs] = x[k];
    }
}

hipError_t launch_zthomas_f64(const ZArgs &a, const ZThomasItem *items, const double *tab, int nitems, hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    if ((!a.solves && (a.nz < 3 || a.nz > 64 * ZTH_CMAX)) || a.items) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zthomas64, dim3((unsigned)nitems), dim3(64), 0, s, a, items, tab);
    return hipGetLastError();
}

}  // namespace pp
