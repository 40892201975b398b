// pp_zmg.hip -- multigrid alternative for the z solve (PP_OPT_ZSOLVER = 2).
//
// The reference keeps a 1D multigrid (MultiGrid::solveMono / solveMulti with red-black
// Gauss-Seidel smoothing, restriction and prolongation, iterating to a residual below 1e-10,
// src/multiGrid.cpp:556-717,1105-1515) beside its Thomas solver; it is not dispatched there.
// Here it is a selectable z solver: one wave per (kx, ky) column, the column and its coarse
// levels in LDS, V-cycles until the max-norm residual drops below tol * max|rhs| or stops
// falling (rounding floor).
//
// Column operator (the reference's z rows, src/triDiag.cpp:1733-1751): u[i-1] + eig u[i] +
// u[i+1] = r[i], first / last diagonal eig - 1 (Dirichlet) or eig + 1 (Neumann).  With
// eig = -2 - k2 this is T - k2 I, T the dz^2-scaled second difference; coarsening by two
// cells (cell-centred) rediscretises it as T - 4^l k2 I with the same end rules, restricts the
// residual as 4 * mean of the two fine cells and prolongs linearly (3/4, 1/4).  A level of at
// most 8 rows, or of odd length, is solved directly by Thomas.  The pinned (0,0) column
// (reference row-0 rule) is solved by the pinned Thomas sweep directly.
#include <hip/hip_runtime.h>

#include "pp_launch.h"
#include "pp_zsolve.h"

namespace pp {

constexpr int MG_LEVELS = 10;

struct MGLevel {
    int n, off;  // rows, offset in the level arrays
    double diag, d0, dn;
};

__device__ inline double fmax_(double a, double b) { return a > b ? a : b; }

__device__ inline double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax_(v, __shfl_xor(v, o));
    return v;
}

// Thomas on one level (lane 0), using w as scratch for the modified super-diagonal
__device__ void mg_direct(const MGLevel &L, double *u, const double *f, double *w)
{
    if (threadIdx.x != 0) return;
    const int n = L.n;
    auto b = [&](int i) { return i == 0 ? L.d0 : (i == n - 1 ? L.dn : L.diag); };
    double bet = b(0);
    u[0] = f[0] / bet;
    for (int i = 1; i < n; i++) {
        w[i] = 1.0 / bet;
        bet = b(i) - w[i];
        u[i] = (f[i] - u[i - 1]) / bet;
    }
    for (int i = n - 2; i >= 0; i--) u[i] -= w[i + 1] * u[i + 1];
}

// one red-black Gauss-Seidel sweep (both colours) on a level, whole wave
__device__ void mg_smooth(const MGLevel &L, double *u, const double *f)
{
    const int n = L.n;
    for (int colour = 0; colour < 2; colour++) {
        for (int i = 2 * (int)threadIdx.x + colour; i < n; i += 2 * (int)blockDim.x) {
            const double b = i == 0 ? L.d0 : (i == n - 1 ? L.dn : L.diag);
            const double um = i > 0 ? u[i - 1] : 0.0, up = i + 1 < n ? u[i + 1] : 0.0;
            u[i] = (f[i] - um - up) / b;
        }
        __syncthreads();
    }
}

__device__ void mg_residual(const MGLevel &L, const double *u, const double *f, double *r)
{
    const int n = L.n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double b = i == 0 ? L.d0 : (i == n - 1 ? L.dn : L.diag);
        const double um = i > 0 ? u[i - 1] : 0.0, up = i + 1 < n ? u[i + 1] : 0.0;
        r[i] = f[i] - (um + b * u[i] + up);
    }
    __syncthreads();
}

// cell-centred linear prolongation of the coarse correction e (length n/2) into u
__device__ void mg_prolong(const MGLevel &F, double *u, const double *e, bool lo_dir, bool hi_dir)
{
    const int nc = F.n / 2;
    for (int I = threadIdx.x; I < nc; I += blockDim.x) {
        const double ec = e[I];
        const double em = I > 0 ? e[I - 1] : (lo_dir ? -ec : ec);
        const double ep = I + 1 < nc ? e[I + 1] : (hi_dir ? -ec : ec);
        u[2 * I] += 0.75 * ec + 0.25 * em;
        u[2 * I + 1] += 0.75 * ec + 0.25 * ep;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_zmg(ZArgs a, int max_cycles, double tol)
{
    extern __shared__ double mg[];
    const int nz = a.nz;
    // level layout: u | f | r, each sum over levels of n_l <= 2 nz
    double *U = mg, *Fv = mg + 2 * nz, *Rv = mg + 4 * nz;
    const int sx = blockIdx.x % a.SX, sy = blockIdx.x / a.SX;
    double *data = reinterpret_cast<double *>(a.data);
    const long zs = (long)a.SX * a.SY, base = (long)sy * a.SX + sx;

    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    bool pin = fabs(eig + 2.) < 1.e-10;
    if (pin && a.zero_mode == 1 && (a.zlo_dirichlet || a.zhi_dirichlet)) pin = false;
    const double k2 = -(eig + 2.0);
    const bool lo_d = a.zlo_dirichlet, hi_d = a.zhi_dirichlet;

    MGLevel lv[MG_LEVELS];
    int nl = 0, off = 0;
    for (int n = nz; nl < MG_LEVELS; n /= 2) {
        const double s = (double)(1L << (2 * nl));  // 4^l
        lv[nl] = {n, off, -2.0 - s * k2, (lo_d ? -3.0 : -1.0) - s * k2, (hi_d ? -3.0 : -1.0) - s * k2};
        nl++;
        off += n;
        if (n <= 8 || (n & 1)) break;
    }
    // the finest level carries the column's own diagonal (no k2 round trip)
    lv[0].diag = eig;
    lv[0].d0 = lo_d ? eig - 1. : eig + 1.;
    lv[0].dn = hi_d ? eig - 1. : eig + 1.;
    if (nl == 1 || pin) {
        // direct solve (short column) or the pinned (0,0) column: Thomas with the row-0 rule
        for (int i = threadIdx.x; i < nz; i += blockDim.x) Fv[i] = data[base + i * zs];
        __syncthreads();
        MGLevel L0 = lv[0];
        L0.diag = eig;
        L0.d0 = lo_d ? eig - 1. : eig + 1.;
        L0.dn = hi_d ? eig - 1. : eig + 1.;
        if (pin) {
            // u0 = r0 (b0 = 1, no super-diagonal), src/triDiag.cpp:1629-1687
            if (threadIdx.x == 0) {
                double *u = U, *w = Rv;
                u[0] = Fv[0];
                double bet = L0.diag;
                w[1] = 0.0;
                u[1] = (Fv[1] - u[0]) / bet;
                for (int i = 2; i < nz; i++) {
                    w[i] = 1.0 / bet;
                    bet = (i == nz - 1 ? L0.dn : L0.diag) - w[i];
                    u[i] = (Fv[i] - u[i - 1]) / bet;
                }
                for (int i = nz - 2; i >= 0; i--) u[i] -= w[i + 1] * u[i + 1];
            }
        } else {
            mg_direct(L0, U, Fv, Rv);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nz; i += blockDim.x) data[base + i * zs] = U[i];
        return;
    }
    double fbig = 0.0;
    for (int i = threadIdx.x; i < nz; i += blockDim.x) {
        const double v = data[base + i * zs];
        Fv[i] = v;
        U[i] = 0.0;
        fbig = fmax_(fbig, fabs(v));
    }
    fbig = wave_max(fbig);
    double rprev = 1.0 / 0.0;
    __syncthreads();
    for (int cycle = 0; cycle < max_cycles; cycle++) {
        // V-cycle: down
        for (int l = 0; l < nl - 1; l++) {
            const MGLevel &L = lv[l];
            double *u = U + L.off, *f = Fv + L.off, *r = Rv + L.off;
            mg_smooth(L, u, f);
            mg_smooth(L, u, f);
            mg_residual(L, u, f, r);
            const MGLevel &C = lv[l + 1];
            for (int I = threadIdx.x; I < C.n; I += blockDim.x) {
                Fv[C.off + I] = 2.0 * (r[2 * I] + r[2 * I + 1]);
                U[C.off + I] = 0.0;
            }
            __syncthreads();
        }
        mg_direct(lv[nl - 1], U + lv[nl - 1].off, Fv + lv[nl - 1].off, Rv + lv[nl - 1].off);
        __syncthreads();
        // up
        for (int l = nl - 2; l >= 0; l--) {
            const MGLevel &L = lv[l];
            mg_prolong(L, U + L.off, U + lv[l + 1].off, lo_d, hi_d);
            mg_smooth(L, U + L.off, Fv + L.off);
            mg_smooth(L, U + L.off, Fv + L.off);
        }
        // convergence: max-norm residual on the finest level below tol * max|f|, or a cycle that
        // reduces it by less than 10% (the rounding floor: a smooth column's residual cancels
        // terms of size |diag| |u|, so its floor is not known in advance)
        mg_residual(lv[0], U, Fv, Rv);
        double rmax = 0.0;
        for (int i = threadIdx.x; i < nz; i += blockDim.x) rmax = fmax_(rmax, fabs(Rv[i]));
        rmax = wave_max(rmax);
        if (rmax <= tol * fbig || rmax > 0.9 * rprev) break;
        rprev = rmax;
    }
    for (int i = threadIdx.x; i < nz; i += blockDim.x) data[base + i * zs] = U[i];
}

hipError_t launch_zsolve_mg_f64(const ZArgs &a, int max_cycles, double tol, hipStream_t s)
{
    const size_t lds = 6 * (size_t)a.nz * sizeof(double);
    if (a.nz < 1 || lds > 64 * 1024 || a.tiled || a.items) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zmg, dim3((unsigned)((long)a.SX * a.SY)), dim3(64), lds, s, a, max_cycles, tol);
    return hipGetLastError();
}

}  // namespace pp
