// pp_verify.hip -- the reference's built-in verification case on the device:
//   pp_init_trigonometric  <- PencilDcmp::initializeTrigonometric (src/pencilDcmp.cpp:1871-2035)
//   pp_get_error           <- PencilDcmp::getError               (src/pencilDcmp.cpp:2040-2208)
//   pp_fill_trigonometric  <- PencilDcmp::fillTrigonometric      (src/pencilDcmp.cpp:5685-5765)
// with the reference's compile-time settings SHIFT = 1 (cell centres), EXACT = 0 and
// COEFF0..2 = 1, 2, 3 (src/include/params.h:36-71).
//
// The reference's manufactured field is complex: its real and imaginary parts are two
// independent real problems of the same operator.  This library solves real fields, so
// pp_init_trigonometric takes the part to set up, and pp_get_error measures the solution of
// that part; the reference's error of the complex field is sqrt(e_re^2 + e_im^2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "pittpack_hip.h"
#include "pp_handle.h"

namespace {

constexpr int kErrBlocks = 1024;  // partial sums of pp_get_error (fixed: a deterministic sum)
constexpr int kErrThreads = 256;

// the case the reference's branches select (src/pencilDcmp.cpp:1982-2012 / 2132-2169)
enum TrigCase { TRIG_XY_PERIODIC = 0, TRIG_ALL_PERIODIC = 1, TRIG_PRODUCT = 2, TRIG_COS2PIZ = 3 };

struct TrigArgs {
    int nx, ny, nz;
    double Xa, Ya, Za, c1, c2, c3;
    int tag[3];
    int tcase, part;
};

// exactValue (src/pencilDcmp.cpp:1837-1866)
__device__ inline double exact_value_dev(double omega, double x, int tg)
{
    if (fabs(omega) < 1.0e-6) return 1.0;
    if (tg == 0 || tg == -1 || tg == -2) return sin(omega * x);
    if (tg == 1) return cos(omega * x);
    return 0.0;
}

// exact solution of the manufactured case at a cell centre, and the matching right-hand side
// (unscaled: the solve multiplies it by dz^2 as assignRhs does)
__device__ inline void trig_point(const TrigArgs &a, size_t p, double &u, double &f)
{
    const double pi = 3.1415926535897932384;
    const double om0 = 1 * pi, om1 = 2 * pi, om2 = 3 * pi;
    const int i = (int)(p % a.nx);
    const int j = (int)((p / a.nx) % a.ny);
    const int k = (int)(p / ((size_t)a.nx * a.ny));
    const double x = a.Xa + i * a.c1 + a.c1 * .5;
    const double y = a.Ya + j * a.c2 + a.c2 * 0.5;
    const double z = a.Za + k * a.c3 + a.c3 * 0.5;
    switch (a.tcase) {
    case TRIG_XY_PERIODIC:
        u = sin(om1 * z) * (a.part ? sin(om1 * (x + y)) : cos(om1 * (x + y)));
        f = -3. * om1 * om1 * u;
        break;
    case TRIG_ALL_PERIODIC:
        u = a.part ? sin(om1 * (x + y + z)) : cos(om1 * (x + y + z));
        f = -3. * om1 * om1 * u;
        break;
    case TRIG_PRODUCT:
        u = a.part ? 0.0
                   : exact_value_dev(om0, x, a.tag[0]) * exact_value_dev(om1, y, a.tag[1]) *
                         exact_value_dev(om2, z, a.tag[2]);
        f = -(om0 * om0 * u + om1 * om1 * u + om2 * om2 * u);
        break;
    default:  // fillTrigonometric's right-hand side and the exact solution getError compares with
        u = cos(2. * pi * z);
        f = -4. * pi * pi * u;
        break;
    }
}

__global__ void k_trig_rhs(TrigArgs a, double *rhs)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        double u, f;
        trig_point(a, p, u, f);
        rhs[p] = f;
    }
}

// per-block sums of (u - u_exact)^2
__global__ __launch_bounds__(kErrThreads) void k_trig_error(TrigArgs a, const double *u, double *partial)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    double s = 0.0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        double ue, f;
        trig_point(a, p, ue, f);
        const double d = u[p] - ue;
        s += d * d;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    __shared__ double w[kErrThreads / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kErrThreads / 64; i++) t += w[i];
        partial[blockIdx.x] = t;
    }
}

TrigArgs trig_args(const pp_solver *h, int part)
{
    TrigArgs a;
    a.nx = h->g.nx;
    a.ny = h->g.ny;
    a.nz = h->g.nz;
    a.Xa = h->g.X[0];
    a.Ya = h->g.X[2];
    a.Za = h->g.X[4];
    a.c1 = h->g.dxyz[0];
    a.c2 = h->g.dxyz[1];
    a.c3 = h->g.dxyz[2];
    for (int d = 0; d < 3; d++) a.tag[d] = h->g.dtag[d];
    const char *bc = h->g.bc;
    if (part < 0) a.tcase = TRIG_COS2PIZ;
    else if (bc[0] == 'P' && bc[2] == 'P' && bc[4] == 'P') a.tcase = TRIG_ALL_PERIODIC;
    else if (bc[0] == 'P' || bc[2] == 'P') a.tcase = TRIG_XY_PERIODIC;
    else a.tcase = TRIG_PRODUCT;
    a.part = part < 0 ? 0 : part;
    return a;
}

unsigned grid_for(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 16384 ? (b ? b : 1) : 16384);
}

}  // namespace

extern "C" {

int pp_init_trigonometric(pp_handle h, int part)
{
    if (!h) return PP_INVALID_HANDLE;
    if (part != 0 && part != 1) return PP_INPUT_ARGS;
    if (!h->bc_set || !h->box_set || !h->coords_set) return PP_CALL_ORDER;
    const TrigArgs a = trig_args(h, part);
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    hipLaunchKernelGGL(k_trig_rhs, dim3(grid_for(n)), dim3(256), 0, h->stream, a, h->d_field);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) return PP_HIP_ERROR;
    h->rhs_set = true;
    h->trig_part = part;
    return PP_SUCCESS;
}

int pp_fill_trigonometric(pp_handle h, double *rhs)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!rhs) return PP_INPUT_ARGS;
    if (!h->box_set || !h->coords_set) return PP_CALL_ORDER;
    const double pi = 3.1415926535897932384;
    const int nx = h->g.nx, ny = h->g.ny, nz = h->g.nz;
    const double Za = h->g.X[4], c3 = h->g.dxyz[2];
    for (int k = 0; k < nz; k++) {
        const double z = Za + k * c3 + c3 * 0.5;
        const double v = -4. * pi * pi * std::cos(2. * pi * z);
        for (size_t p = 0; p < (size_t)nx * ny; p++) rhs[(size_t)k * nx * ny + p] = v;
    }
    return PP_SUCCESS;
}

int pp_get_error(pp_handle h, double *err)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!err) return PP_INPUT_ARGS;
    if (!h->bc_set || !h->box_set || !h->coords_set) return PP_CALL_ORDER;
    if (!h->d_partials && hipMalloc((void **)&h->d_partials, kErrBlocks * sizeof(double)) != hipSuccess) {
        h->d_partials = nullptr;
        return PP_ALLOCATION_FAIL;
    }
    const TrigArgs a = trig_args(h, h->trig_part);
    hipLaunchKernelGGL(k_trig_error, dim3(kErrBlocks), dim3(kErrThreads), 0, h->stream, a, h->d_field,
                       h->d_partials);
    double part[kErrBlocks];
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(part, h->d_partials, sizeof(part), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return PP_HIP_ERROR;
    double s = 0.0;
    for (int b = 0; b < kErrBlocks; b++) s += part[b];
    *err = std::sqrt(s / ((double)a.nx * a.ny * a.nz));
    return PP_SUCCESS;
}

}  // extern "C"
