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
// the Sherman-Morrison correction), restated serially in FP64 by k_smt_v1.
//
// Lengths without a compiled complex plan take three passes instead of one: real z-line
// transforms of any length (the line kernels) around a division kernel (k_zper_divide).
#include <hip/hip_runtime.h>

#include "pp_fft.h"
#include "pp_launch.h"
#include "pp_line.h"
#include "pp_zsolve.h"

namespace pp {

// ZPerArgs (pp_launch.h): data [nz][SY][SX] in place; tw = exp(-2 pi i e / nz), e < nz,
// interleaved; c2 = 2 cos(2 pi k / nz), k < nz

template <int M, int T, int LINES>
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
    const int sx0 = blockIdx.x * LINES + l;
    const int sy = blockIdx.y;
    const bool valid = sx0 < a.SX;
    const int sx = valid ? sx0 : a.SX - 1;
    const long base = (long)sy * a.SX + sx;
    const long stride = (long)a.SX * a.SY;

    cpx<R> v[E];
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = {valid ? a.data[base + (long)(t + T * m) * stride] : 0.0, 0.0};
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
    if (valid && !singular) {
#pragma unroll
        for (int m = 0; m < E; m++) a.data[base + (long)(t + T * m) * stride] = v[m].x;
    }
}

// the reference's singular periodic branch on one column (sx, sy), FP64, serial
__global__ void k_smt_v1(double *data, int SX, int SY, int nz, int sx, int sy, double diag)
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

template <int M>
static hipError_t launch_zper_m(const ZPerArgs &a, hipStream_t s)
{
    constexpr int T = line_threads(M, 0);
    // 16 columns (128-byte rows) while the block's line slices fit in 80 KB, else 8
    constexpr int LINES = (16 * line_slots<M>() * 16 <= 80 * 1024) ? 16 : 8;
    const size_t lds = (size_t)LINES * line_slots<M>() * sizeof(cpx<double>);
    auto kern = k_zper<M, T, LINES>;
    if (lds > 64 * 1024) {
        static bool attr = false;
        if (!attr) {
            hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr = true;
        }
    }
    dim3 grid((unsigned)((a.SX + LINES - 1) / LINES), (unsigned)a.SY);
    hipLaunchKernelGGL(kern, grid, dim3(T * LINES), lds, s, a);
    return hipGetLastError();
}

bool zperiodic_supported(int nz) { return nz >= 8 && line_threads(nz, 0) > 0; }

// Periodic z of any other length: the field's z lines go through the real line transforms
// (R2HC forward, HC2R inverse, pp_line.h / pp_line_bs.h, launched by the host on the
// [nz][SY*SX] view) and this kernel divides half-complex entry p of column (sx, sy) by
// nz * (eig + 2 cos(2 pi p / nz)) -- modes k and nz - k share the factor, so the real and
// imaginary parts of a mode divide alike.  The singular column only takes 1 / nz, so the
// inverse transform restores its right-hand side for the reference's branch (k_smt_v1).
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
    a.data[i] = a.data[i] / den;
}

hipError_t launch_zper_divide(const ZPerArgs &a, hipStream_t s)
{
    const long n = (long)a.SX * a.SY * a.nz;
    hipLaunchKernelGGL(k_zper_divide, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zperiodic_f64(const ZPerArgs &a, hipStream_t s)
{
    switch (a.nz) {
#define PPM(m) \
    case m: return launch_zper_m<m>(a, s);
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_smt_v1(double *data, int SX, int SY, int nz, int sx, int sy, double diag, hipStream_t s)
{
    hipLaunchKernelGGL(k_smt_v1, dim3(1), dim3(256), 3 * (size_t)nz * sizeof(double), s, data, SX, SY, nz, sx, sy,
                       diag);
    return hipGetLastError();
}

}  // namespace pp
