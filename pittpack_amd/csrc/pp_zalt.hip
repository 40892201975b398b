// pp_zalt.hip -- the reference's alternative tridiagonal solvers as selectable z solvers:
// parallel cyclic reduction (PP_OPT_ZSOLVER = 3) and cyclic reduction (PP_OPT_ZSOLVER = 4).
//
// Both work on the column's rows normalised by their diagonal (diag 1, off-diagonals and
// right-hand side divided by it), as the reference's drivers do (solvePCR / solveCRP and
// their helpers fillInArrayContigNormalize, offDiagNormalize, imposeBoundaryon*,
// src/pencilDcmp.cpp:5239-5488).  The column operator is the one every other z solver here
// uses: u[i-1] + eig u[i] + u[i+1], first / last diagonal eig - 1 (Dirichlet ghost cell) or
// eig + 1 (Neumann), src/triDiag.cpp:1733-1751.
//
//  * PCR (TriDiag::pcr, src/triDiag.cpp:1442-1542): every row eliminates its neighbours at
//    distance s = 1, 2, 4, ... (r = 1 / (1 - a_i c_{i-s} - c_i a_{i+s}); a_i <- -r a_i a_{i-s};
//    c_i <- -r c_i c_{i+s}; d_i <- r (d_i - a_i d_{i-s} - c_i d_{i+s})) until no row couples
//    to another; then u_i = d_i.
//  * CR (TriDiag::crp, src/triDiag.cpp:1285-1437): odd-even reduction -- at half-stride h
//    the rows 2h j + 2h - 1 eliminate their neighbours at distance h (same update, diagonal
//    renormalised), down to the last row standing, then back substitution
//    u_i = d_i - a_i u_{i-h} - c_i u_{i+h}.  The reference's version takes uniform
//    coefficients and n = 2^k - 1 only, without the ghost-cell end rows; here the end rows
//    are carried per row and the column is padded to 2^k - 1 rows with identity rows.
//
// One 64-lane wave per (kx, ky) column, its rows in LDS.  The pinned (0,0) column (the
// reference's row-0 rule, src/triDiag.cpp:1629-1687) is solved by the pinned Thomas sweep,
// which neither reduction applies to (the reference's PCR/CRP drivers have no such branch).
#include <hip/hip_runtime.h>

#include "pp_launch.h"
#include "pp_zsolve.h"

namespace pp {

struct ColumnRows {
    double eig, b0, bn;
    bool pin;
};

__device__ inline ColumnRows column_rows(const ZArgs &a, int sx, int sy)
{
    ColumnRows c;
    c.eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    c.pin = fabs(c.eig + 2.) < 1.e-10;
    if (c.pin && a.zero_mode == 1 && (a.zlo_dirichlet || a.zhi_dirichlet)) c.pin = false;
    c.b0 = a.zlo_dirichlet ? c.eig - 1. : c.eig + 1.;
    c.bn = a.zhi_dirichlet ? c.eig - 1. : c.eig + 1.;
    return c;
}

// pinned column: u0 = r0, then Thomas on rows 1.. (lane 0; w: scratch of n doubles)
__device__ void pinned_thomas(double *u, double *w, int n, double eig, double bn)
{
    if (threadIdx.x != 0) return;
    double bet = eig;
    w[1] = 0.0;
    u[1] = (u[1] - u[0]) / bet;
    for (int i = 2; i < n; i++) {
        w[i] = 1.0 / bet;
        bet = (i == n - 1 ? bn : eig) - w[i];
        u[i] = (u[i] - u[i - 1]) / bet;
    }
    for (int i = n - 2; i >= 1; i--) u[i] -= w[i + 1] * u[i + 1];
}

// normalised row i of the column: off-diagonals and right-hand side divided by the diagonal
__device__ inline void norm_row(const ColumnRows &c, int i, int n, double d, double &ai, double &ci, double &di)
{
    const double b = i == 0 ? c.b0 : (i == n - 1 ? c.bn : c.eig);
    const double inv = 1.0 / b;
    ai = i == 0 ? 0.0 : inv;
    ci = i == n - 1 ? 0.0 : inv;
    di = d * inv;
}

constexpr int ALT_ROWS = 16;  // rows per lane kept in registers during a PCR step (nz <= 1024)

__global__ void __launch_bounds__(64) k_zpcr(ZArgs a)
{
    extern __shared__ double zp[];
    const int n = a.nz;
    double *A = zp, *C = zp + n, *D = zp + 2 * n;
    const int sx = blockIdx.x % a.SX, sy = blockIdx.x / a.SX;
    double *data = reinterpret_cast<double *>(a.data);
    const long zs = (long)a.SX * a.SY, base = (long)sy * a.SX + sx;
    const ColumnRows c = column_rows(a, sx, sy);
    if (c.pin) {
        for (int i = threadIdx.x; i < n; i += 64) D[i] = data[base + i * zs];
        __syncthreads();
        pinned_thomas(D, A, n, c.eig, c.bn);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 64) data[base + i * zs] = D[i];
        return;
    }
    for (int i = threadIdx.x; i < n; i += 64) norm_row(c, i, n, data[base + i * zs], A[i], C[i], D[i]);
    __syncthreads();
    for (int s = 1; s < n; s <<= 1) {
        // every row from the previous level's values: read into registers, then write
        double na[ALT_ROWS], nc[ALT_ROWS], nd[ALT_ROWS];
#pragma unroll
        for (int k = 0; k < ALT_ROWS; k++) {
            const int i = threadIdx.x + 64 * k;
            if (i < n) {
                const double ai = A[i], ci = C[i], di = D[i];
                const bool lo = i - s >= 0, hi = i + s < n;
                const double al = lo ? A[i - s] : 0.0, cl = lo ? C[i - s] : 0.0, dl = lo ? D[i - s] : 0.0;
                const double ah = hi ? A[i + s] : 0.0, ch = hi ? C[i + s] : 0.0, dh = hi ? D[i + s] : 0.0;
                const double r = 1.0 / (1.0 - ai * cl - ci * ah);
                na[k] = -r * ai * al;
                nc[k] = -r * ci * ch;
                nd[k] = r * (di - ai * dl - ci * dh);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < ALT_ROWS; k++) {
            const int i = threadIdx.x + 64 * k;
            if (i < n) {
                A[i] = na[k];
                C[i] = nc[k];
                D[i] = nd[k];
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += 64) data[base + i * zs] = D[i];
}

__global__ void __launch_bounds__(64) k_zcr(ZArgs a, int np)
{
    // np = 2^k - 1 >= nz rows; rows from nz on are identity rows (a = c = 0, d = 0)
    extern __shared__ double zc[];
    const int n = a.nz;
    double *A = zc, *C = zc + np, *D = zc + 2 * np;
    const int sx = blockIdx.x % a.SX, sy = blockIdx.x / a.SX;
    double *data = reinterpret_cast<double *>(a.data);
    const long zs = (long)a.SX * a.SY, base = (long)sy * a.SX + sx;
    const ColumnRows c = column_rows(a, sx, sy);
    if (c.pin) {
        for (int i = threadIdx.x; i < n; i += 64) D[i] = data[base + i * zs];
        __syncthreads();
        pinned_thomas(D, A, n, c.eig, c.bn);
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += 64) data[base + i * zs] = D[i];
        return;
    }
    for (int i = threadIdx.x; i < np; i += 64) {
        if (i < n) norm_row(c, i, n, data[base + i * zs], A[i], C[i], D[i]);
        else A[i] = C[i] = D[i] = 0.0;
    }
    __syncthreads();
    // forward reduction: at half-stride h the rows 2h j + 2h - 1 absorb rows at distance h
    int h = 1;
    for (; 2 * h <= np; h <<= 1) {
        const int m = (np + 1) / (2 * h) - 1;  // rows reduced at this level
        for (int j = threadIdx.x; j < m; j += 64) {
            const int i = 2 * h * j + 2 * h - 1;
            const double ai = A[i], ci = C[i];
            const double b = 1.0 - ai * C[i - h] - ci * A[i + h];
            const double inv = 1.0 / b;
            const double nd = (D[i] - ai * D[i - h] - ci * D[i + h]) * inv;
            A[i] = -ai * A[i - h] * inv;
            C[i] = -ci * C[i + h] * inv;
            D[i] = nd;
        }
        __syncthreads();
    }
    // the middle row (np - 1) / 2 no longer couples to any row: u = d.  Back substitution:
    // at half-stride h the rows 2h j + h - 1 take their neighbours at distance h
    for (h >>= 1; h >= 1; h >>= 1) {
        const int m = (np + 1) / (2 * h);  // rows solved at this level
        for (int j = threadIdx.x; j < m; j += 64) {
            const int i = 2 * h * j + h - 1;
            const double ul = i - h >= 0 ? D[i - h] : 0.0, uh = i + h < np ? D[i + h] : 0.0;
            D[i] = D[i] - A[i] * ul - C[i] * uh;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += 64) data[base + i * zs] = D[i];
}

bool zsolve_alt_supported(int nz) { return nz >= 4 && nz <= 64 * ALT_ROWS; }

hipError_t launch_zsolve_pcr_full_f64(const ZArgs &a, hipStream_t s)
{
    if (!zsolve_alt_supported(a.nz) || a.tiled || a.items) return hipErrorInvalidValue;
    const size_t lds = 3 * (size_t)a.nz * sizeof(double);
    hipLaunchKernelGGL(k_zpcr, dim3((unsigned)((long)a.SX * a.SY)), dim3(64), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_zsolve_cr_f64(const ZArgs &a, hipStream_t s)
{
    if (!zsolve_alt_supported(a.nz) || a.tiled || a.items) return hipErrorInvalidValue;
    int np = 1;
    while (np < a.nz) np = 2 * np + 1;
    const size_t lds = 3 * (size_t)np * sizeof(double);
    if (lds > 64 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zcr, dim3((unsigned)((long)a.SX * a.SY)), dim3(64), lds, s, a, np);
    return hipGetLastError();
}

}  // namespace pp
