// pp_zperiodic.hip -- z solve for periodic z faces.
//
// The reference solves each (kx, ky) column's cyclic tridiagonal system (diagonal eig, unit
// off-diagonals, wrap-around corners) by Sherman-Morrison on top of two Thomas sweeps
// (shermanMorrisonThomas, src/triDiag.cpp:1961-2108).  That system is circulant, so here a
// block diagonalises it exactly instead: complex FFT of the column along z, division of mode
// k by eig + 2 cos(2 pi k / nz), inverse FFT -- one read and one write of the column, with the
// Stockham passes of the line transforms (pp_fft.h) running on z-columns.
//
// The singular column (|eig + 2| < 1e-10: x and y in {P, N}, mode (0, 0)) follows the
// reference's own branch (shermanMorrisonThomasV1, src/triDiag.cpp:2112-2255: first row
// replaced by the average of the first and last right-hand sides, pinned Thomas sweeps and
// the Sherman-Morrison correction), restated serially in FP64 by k_smt_v1.  The mixed form
// solves that column from FP64 plane sums instead (k_mode00, pp_mixed.hip) and has it
// written as zeros here.
//
// Lengths without a compiled complex plan take three passes instead of one: real z-line
// transforms of any length (the line kernels) around a division kernel (k_zper_divide).
//
// Storage is FP64 (FP64 form) or FP32 (the mixed form's work field); the arithmetic of the
// one-pass kernel and of the division is FP64 in both.  Batched launches take one block per
// item (solve, column block, y row) over the solves' descriptors.
#include <hip/hip_runtime.h>

#include "pp_fft.h"
#include "pp_launch.h"
#include "pp_line.h"
#include "pp_zsolve.h"

namespace pp {

// ZPerArgs (pp_launch.h): data [nz][SY][SX] in place; tw = exp(-2 pi i e / nz), e < nz,
// interleaved; c2 = 2 cos(2 pi k / nz), k < nz.  S is the storage type: the FP64 form and the
// mixed form's FP32 work field both compute in FP64 (the transform and the division by the
// mode's eigenvalue), so near-singular columns need no separate treatment in either form.
template <typename S, int M, int T, int LINES>
__global__ void __launch_bounds__(T *LINES) k_zper(ZPerArgs a)
{
    using R = double;
    constexpr int E = M / T;
    constexpr int LS = line_slots<M>();
    extern __shared__ __align__(16) unsigned char smem_zp[];
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_zp);
    const int tid = threadIdx.x;
    const int l = tid % LINES, t = tid / LINES;
    cpx<R> *L = lds + l * LS;
    int bx = blockIdx.x, sy = blockIdx.y;
    S *data = reinterpret_cast<S *>(a.data);
    if (a.items) {
        // batched: this block's solve, its sizes and eigenvalue factors
        const ZItem it = a.items[blockIdx.x];
        const ZSolve sv = a.solves[it.solve];
        data += sv.offset;
        a.SX = sv.SX;
        a.SY = sv.SY;
        a.Cx = sv.Cx;
        a.Cy = sv.Cy;
        a.lamx = sv.lamx;
        a.lamy = sv.lamy;
        bx = it.bx;
        sy = it.by;
    }
    const int sx0 = bx * LINES + l;
    const bool valid = sx0 < a.SX;
    const int sx = valid ? sx0 : a.SX - 1;
    const long base = (long)sy * a.SX + sx;
    const long stride = (long)a.SX * a.SY;

    cpx<R> v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = {valid ? (R)data[base + (long)(t + T * m) * stride] : 0.0, 0.0};
    FFTSeeds<R, M, T> fs;
    fs.load(t, a.tw);
    // eigenvalue of this (kx, ky) column, src/pencilDcmp.cpp:4393-4417, without contraction
    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    const bool singular = fabs(eig + 2.) < 1.e-10;

    line_fft<R, M, T, -1>(v, L, t, fs);
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int k = t + T * m;
        const double den = (eig + a.c2[k]) * (double)M;
        const double s = singular ? 0.0 : 1.0 / den;
        v[m] = c_scale(v[m], s);
    }
    line_fft<R, M, T, +1>(v, L, t, fs);
    // the singular column: zeros (pinned_zero, v is zero) or left for k_smt_v1
    if (valid && (!singular || a.pinned_zero)) {
#pragma unroll
        for (int m = 0; m < E; m++) data[base + (long)(t + T * m) * stride] = (S)v[m].x;
    }
}

// the reference's singular periodic branch on one column (sx, sy), FP64, serial
__device__ void smt_v1_body(double *data, int SX, int SY, int nz, int sx, int sy, double diag)
{
    extern __shared__ double sv[];
    double *rh = sv, *rh1 = sv + nz, *tmp = sv + 2 * nz;
    const long stride = (long)SX * SY, base = (long)sy * SX + sx;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) rh[i] = data[base + i * stride];
    __syncthreads();
    if (threadIdx.x == 0) {
        const int N = nz;
        const double alpha = 1.0, beta = 1.0;
        // shermanMorrisonThomasV1 with b = {1, diag, diag}
        const double b0 = 1.0;
        const double gamma = -b0;
        const double bb[3] = {b0 - gamma, diag, diag - alpha * beta / gamma};
        // thomasLowMemNoBCV1 (src/triDiag.cpp:1629-1687): row 0 decoupled
        auto sweep = [&](double *r) {
            double bet;
            r[0] = r[0] / (bet = bb[0]);
            int j = 1;
            tmp[j] = 0.0;
            bet = bb[1] - tmp[j];
            r[1] = (r[1] - r[j - 1]) / bet;
            for (j = 2; j < N - 1; j++) {
                tmp[j] = 1.0 / bet;
                bet = bb[1] - tmp[j];
                r[j] = (r[j] - r[j - 1]) / bet;
            }
            j = N - 1;
            tmp[j] = 1.0 / bet;
            bet = bb[2] - tmp[j];
            r[j] = (r[j] - r[j - 1]) / bet;
            for (j = N - 2; j >= 0; j--) r[j] -= tmp[j + 1] * r[j + 1];
        };
        rh[0] = (rh[0] + rh[N - 1]) * 0.5;
        sweep(rh);
        rh1[0] = gamma;
        rh1[N - 1] = alpha;
        for (int i = 1; i < N - 1; i++) rh1[i] = 0.0;
        sweep(rh1);
        const double part1 = (rh[0] + beta * rh[N - 1] / gamma);
        const double part2 = (1. + rh1[0] + beta * rh1[N - 1] / gamma);
        const double fact = part1 / part2;
        for (int i = 0; i < N; i++) rh[i] -= fact * rh1[i];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nz; i += blockDim.x) data[base + i * stride] = rh[i];
}

__global__ void k_smt_v1(double *data, int SX, int SY, int nz, int sx, int sy, double diag)
{
    smt_v1_body(data, SX, SY, nz, sx, sy, diag);
}

// batched: block b solves the singular column of items[b]
__global__ void k_smt_v1_batch(double *data, const SmtItem *items)
{
    const SmtItem it = items[blockIdx.x];
    smt_v1_body(data + it.offset, it.SX, it.SY, it.nz, it.sx, it.sy, it.diag);
}

template <int M>
constexpr int zper_lines_m()
{
    // 16 columns (128-byte FP64 rows) while the block's line slices fit in 80 KB, else 8
    return (16 * line_slots<M>() * 16 <= 80 * 1024) ? 16 : 8;
}

// grid: column blocks x SY (one solve), or nitems blocks (batched, a.items set)
template <typename S, int M>
static hipError_t launch_zper_m(const ZPerArgs &a, int nitems, hipStream_t s)
{
    constexpr int T = line_threads(M, 0);
    constexpr int LINES = zper_lines_m<M>();
    const size_t lds = (size_t)LINES * line_slots<M>() * sizeof(cpx<double>);
    auto kern = k_zper<S, M, T, LINES>;
    if (lds > 64 * 1024) {
        static bool attr = false;
        if (!attr) {
            hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr = true;
        }
    }
    const dim3 grid = a.items ? dim3((unsigned)nitems) : dim3((unsigned)((a.SX + LINES - 1) / LINES), (unsigned)a.SY);
    if (a.items && nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(kern, grid, dim3(T * LINES), lds, s, a);
    return hipGetLastError();
}

bool zperiodic_supported(int nz) { return nz >= 8 && line_threads(nz, 0) > 0; }

int zper_lines(int nz)
{
    switch (nz) {
#define PPM(m) \
    case m: return zper_lines_m<m>();
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return 0;
    }
}

// Periodic z of any other length: the field's z lines go through the real line transforms
// (R2HC forward, HC2R inverse, pp_line.h / pp_line_bs.h, launched by the host on the
// [nz][SY*SX] view) and this kernel divides half-complex entry p of column (sx, sy) by
// nz * (eig + 2 cos(2 pi p / nz)) -- modes k and nz - k share the factor, so the real and
// imaginary parts of a mode divide alike.  The singular column only takes 1 / nz, so the
// inverse transform restores its right-hand side for the reference's branch (k_smt_v1), or
// is zeroed (pinned_zero).
template <typename S>
__global__ void k_zper_divide(ZPerArgs a)
{
    const long plane = (long)a.SX * a.SY;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= plane * a.nz) return;
    const int p = (int)(i / plane);
    const long c = i - (long)p * plane;
    const int sy = (int)(c / a.SX), sx = (int)(c - (long)sy * a.SX);
    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    const bool singular = fabs(eig + 2.) < 1.e-10;
    const double den = singular ? (double)a.nz : (eig + a.c2[p]) * (double)a.nz;
    S *data = reinterpret_cast<S *>(a.data);
    data[i] = (singular && a.pinned_zero) ? (S)0 : (S)((double)data[i] / den);
}

template <typename S>
static hipError_t launch_zper_divide_s(const ZPerArgs &a, hipStream_t s)
{
    const long n = (long)a.SX * a.SY * a.nz;
    hipLaunchKernelGGL(k_zper_divide<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zper_divide(const ZPerArgs &a, hipStream_t s) { return launch_zper_divide_s<double>(a, s); }

hipError_t launch_zper_divide_f32(const ZPerArgs &a, hipStream_t s) { return launch_zper_divide_s<float>(a, s); }

template <typename S>
static hipError_t launch_zperiodic_s(const ZPerArgs &a, int nitems, hipStream_t s)
{
    switch (a.nz) {
#define PPM(m) \
    case m: return launch_zper_m<S, m>(a, nitems, s);
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zperiodic_f64(const ZPerArgs &a, hipStream_t s) { return launch_zperiodic_s<double>(a, 0, s); }

hipError_t launch_zperiodic_f32(const ZPerArgs &a, hipStream_t s) { return launch_zperiodic_s<float>(a, 0, s); }

hipError_t launch_zperiodic_f64_batch(const ZPerArgs &a, int nitems, hipStream_t s)
{
    return a.items ? launch_zperiodic_s<double>(a, nitems, s) : hipErrorInvalidValue;
}

hipError_t launch_zperiodic_f32_batch(const ZPerArgs &a, int nitems, hipStream_t s)
{
    return a.items ? launch_zperiodic_s<float>(a, nitems, s) : hipErrorInvalidValue;
}

hipError_t launch_smt_v1(double *data, int SX, int SY, int nz, int sx, int sy, double diag, hipStream_t s)
{
    hipLaunchKernelGGL(k_smt_v1, dim3(1), dim3(256), 3 * (size_t)nz * sizeof(double), s, data, SX, SY, nz, sx, sy,
                       diag);
    return hipGetLastError();
}

hipError_t launch_smt_v1_batch(double *data, const SmtItem *items, int nitems, int nz_max, hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(k_smt_v1_batch, dim3((unsigned)nitems), dim3(256), 3 * (size_t)nz_max * sizeof(double), s, data,
                       items);
    return hipGetLastError();
}

}  // namespace pp
