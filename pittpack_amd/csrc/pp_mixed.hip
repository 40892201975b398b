// pp_mixed.hip -- kernels of the mixed-precision form: FP32 transforms and z solves, with
// FP64 where FP32 cannot deliver FP64 accuracy after refinement.
//
//  * k_zfix64    : FP64 Thomas for the near-singular (kx, ky) columns (|eig+2| < tau); the
//                  FP32 z solve leaves those columns untouched.  Same algebra as
//                  thomasLowMem without the pinning branch (src/triDiag.cpp:1693-1822).
//  * k_mode00    : exact FP64 path of the pinned (0,0) column (x and y periodic or Neumann,
//                  z Neumann-Neumann, or the reference row-0 rule): the column's right-hand
//                  side is the xy-sum of the effective RHS per plane, the pinned system
//                  (row 0: u0 = r0, src/triDiag.cpp:1629-1687) is solved in FP64, and its
//                  physical contribution -- constant over each plane -- is added by the final
//                  inverse-x store.
//  * k_residual  : r = (rhs*dz^2 - Dirichlet term) - L u, FP64, for the iterative
//                  refinement; L is the dz^2-scaled 7-point operator with the ghost-cell
//                  closures the spectral solve diagonalises (P wrap, N mirror, D antisym).
#include <hip/hip_runtime.h>

#include "pp_line.h"
#include "pp_mixed.h"
#include "pp_zsolve.h"

namespace pp {

struct Row3d {
    double a, b, c;
};

// non-pinned column coefficients of global row i
__device__ __forceinline__ Row3d zrow(int i, int n, double eig, double b0, double bn)
{
    if (i == 0) return {0.0, b0, 1.0};
    if (i == n - 1) return {1.0, bn, 0.0};
    return {1.0, eig, 1.0};
}

// cols[b] = {sx, sy, member, 0}; batched launches (a.solves set) take the member's fields
template <typename R>
__global__ void __launch_bounds__(256) k_zfix64(ZArgs a0, const int4 *cols)
{
    extern __shared__ double zf[];
    const int4 sc = cols[blockIdx.x];
    ZArgs a = a0;
    if (a0.solves) {
        const ZSolve sv = a0.solves[sc.z];
        a.data = reinterpret_cast<R *>(a0.data) + sv.offset;
        a.SX = sv.SX;
        a.SY = sv.SY;
        a.nz = sv.nz;
        a.zlo_dirichlet = sv.zlo_dirichlet;
        a.zhi_dirichlet = sv.zhi_dirichlet;
        a.Cx = sv.Cx;
        a.Cy = sv.Cy;
        a.lamx = sv.lamx;
        a.lamy = sv.lamy;
    }
    double *col = zf;           // nz values
    double *cp = zf + a.nz;     // nz values
    const long zs = (long)a.SX * a.SY;
    const long off = (long)sc.y * a.SX + sc.x;
    R *data = reinterpret_cast<R *>(a.data);
    const int n = a.nz;
    for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = (double)data[off + (long)i * zs];
    __syncthreads();
    if (threadIdx.x == 0) {
        const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sc.x], a.Cx)), __dmul_rn(a.lamy[sc.y], a.Cy));
        const double b0 = a.zlo_dirichlet ? eig - 1. : eig + 1.;
        const double bn = a.zhi_dirichlet ? eig - 1. : eig + 1.;
        Row3d r = zrow(0, n, eig, b0, bn);
        double inv = 1.0 / r.b;
        cp[0] = r.c * inv;
        col[0] = col[0] * inv;
        for (int i = 1; i < n; i++) {
            r = zrow(i, n, eig, b0, bn);
            inv = 1.0 / (r.b - r.a * cp[i - 1]);
            cp[i] = r.c * inv;
            col[i] = (col[i] - r.a * col[i - 1]) * inv;
        }
        for (int i = n - 2; i >= 0; i--) col[i] = col[i] - cp[i] * col[i + 1];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) data[off + (long)i * zs] = (R)col[i];
}

// One block, FP64, for the pinned (0,0) column.  R(z) = sum_y line_sums[y + ny z] is the
// column's right-hand side (plane sums of the forward-x input).
//  initial (correct == 0): solve P u = R with P the pinned operator (row 0: u0 = r0,
//      interior 1 -2 1, last row bn); plane_add[z] = u(z)/(NxNy);
//      plane_add[nz] = (b0 u0 + u1 - S0)/(NxNy), the row-0 defect of the unpinned operator,
//      and plane_add[nz+1] = S0 = R(0).
//  correction (correct == 1): the refinement residual's plane sums give rows z >= 1 of
//      P du = R; row 0 is S0 - (plane-0 sum of the current u), assembled from the residual
//      kernel's per-block partial sums upart[0..nparts).  plane_add[z] = du(z)/(NxNy).
__device__ void mode00_body(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                            double *plane_add, const double *upart, int nparts, int correct)
{
    extern __shared__ double m0[];
    double *r = m0;
    double *cp = m0 + nz;
    __shared__ double red[1024];
    for (int z = threadIdx.x; z < nz; z += blockDim.x) {
        double acc = 0.0;
        for (int y = 0; y < ny; y++) acc += line_sums[(long)z * ny + y];
        r[z] = acc;
    }
    double part = 0.0;
    if (correct)
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) part += upart[i];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double s0 = correct ? plane_add[nz + 1] : r[0];
        if (correct) r[0] = s0 - red[0];
        cp[0] = 0.0;
        for (int i = 1; i < nz; i++) {
            const double b = (i == nz - 1) ? bn : -2.0;
            const double c = (i == nz - 1) ? 0.0 : 1.0;
            const double inv = 1.0 / (b - cp[i - 1]);
            cp[i] = c * inv;
            r[i] = (r[i] - r[i - 1]) * inv;
        }
        for (int i = nz - 2; i >= 0; i--) r[i] = r[i] - cp[i] * r[i + 1];
        if (!correct) {
            plane_add[nz] = (b0 * r[0] + r[1] - s0) * inv_scale;
            plane_add[nz + 1] = s0;
        }
    }
    __syncthreads();
    for (int z = threadIdx.x; z < nz; z += blockDim.x) plane_add[z] = r[z] * inv_scale;
}

__global__ void __launch_bounds__(1024) k_mode00(const double *line_sums, int ny, int nz, double b0, double bn,
                                                 double inv_scale, double *plane_add, const double *upart, int nparts,
                                                 int correct)
{
    mode00_body(line_sums, ny, nz, b0, bn, inv_scale, plane_add, upart, nparts, correct);
}

// batched: block b handles pinned member items[b]
__global__ void __launch_bounds__(1024) k_mode00_batch(const Mode00Item *items, int correct)
{
    const Mode00Item it = items[blockIdx.x];
    mode00_body(it.line_sums, it.ny, it.nz, it.b0, it.bn, it.inv_scale, it.plane_add, it.upart, it.nparts, correct);
}

constexpr int RES_ZC = 32;  // planes marched per thread

// batched launches: block b works on items[b] = {member, bx, by, bz} with solves[member]
template <typename TOut>
__global__ void __launch_bounds__(256) k_residual(ResArgs a0, const ResArgs *solves, const int4 *items)
{
    ResArgs a = a0;
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z, gx = gridDim.x;
    if (items) {
        const int4 it = items[blockIdx.x];
        a = solves[it.x];
        bx = it.y;
        by = it.z;
        bz = it.w;
        gx = (a.nx + 63) / 64;
    }
    const int x = bx * 64 + (threadIdx.x & 63);
    const int y = by * 4 + (threadIdx.x >> 6);
    const int z0 = bz * RES_ZC;
    const bool valid = x < a.nx && y < a.ny;
    const long sy = a.nx, sz = (long)a.nx * a.ny;
    const double *u = a.u;
    if (a.upart && bz == 0) {
        // plane-0 sum of u, one partial per (x, y) block
        __shared__ double red[256];
        red[threadIdx.x] = valid ? u[x + sy * y] : 0.0;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.upart[by * gx + bx] = red[0];
    }
    if (!valid) return;
    TOut *r = reinterpret_cast<TOut *>(a.r);
    const double *ex = a.dtab, *ey = a.dtab + a.nx, *ez = a.dtab + a.nx + a.ny;
    const double *ef = a.dtab + a.nx + a.ny + a.nz;
    // neighbour of a boundary cell through the ghost closure the spectral solve
    // diagonalises: P wraps, N mirrors (ghost = inside), D is antisymmetric (ghost = -inside);
    // as an index and a sign per direction, fixed for this thread's (x, y)
    auto nb = [](char bc, int i, int wrap, long step, long &off, double &sg) {
        if (bc == 'P') {
            off = (long)(wrap - i) * step;
            sg = 1.0;
        } else {
            off = 0;
            sg = (bc == 'N') ? 1.0 : -1.0;
        }
    };
    long oxm = -1, oxp = 1, oym = -sy, oyp = sy;
    double gxm = 1.0, gxp = 1.0, gym = 1.0, gyp = 1.0;
    if (x == 0) nb(a.bc[0], x, a.nx - 1, 1, oxm, gxm);
    if (x == a.nx - 1) nb(a.bc[1], x, 0, 1, oxp, gxp);
    if (y == 0) nb(a.bc[2], y, a.ny - 1, sy, oym, gym);
    if (y == a.ny - 1) nb(a.bc[3], y, 0, sy, oyp, gyp);
    // Dirichlet terms of the effective right-hand side: x and y faces are a per-(x, y)
    // constant times ez[z], z faces a constant on the first and last plane
    double cxy = 0.0, czlo = 0.0, czhi = 0.0;
    if (a.dmask) {
        if ((a.dmask & 1) && x == 0) cxy += 2. * ef[0] * ey[y];
        if ((a.dmask & 2) && x == a.nx - 1) cxy += 2. * ef[1] * ey[y];
        if ((a.dmask & 4) && y == 0) cxy += 2. * ex[x] * ef[2];
        if ((a.dmask & 8) && y == a.ny - 1) cxy += 2. * ex[x] * ef[3];
        if (a.dmask & 16) czlo = 2. * ex[x] * ey[y] * ef[4];
        if (a.dmask & 32) czhi = 2 * ex[x] * ey[y] * ef[5];
    }
    const double defect = a.pin_defect ? a.pin_defect[0] : 0.0;
    const double *uc_p = u + x + sy * y;
    const double *f_p = a.f + x + sy * y;
    TOut *r_p = r + x + sy * y;
    double uc = uc_p[sz * z0];
    double um;
    if (z0 > 0) um = uc_p[sz * (z0 - 1)];
    else um = a.bc[4] == 'P' ? uc_p[sz * (a.nz - 1)] : (a.bc[4] == 'N' ? uc : -uc);
    auto plane = [&](int z) {
        const long o = sz * z;
        double up;
        if (z + 1 < a.nz) up = uc_p[o + sz];
        else up = a.bc[5] == 'P' ? uc_p[0] : (a.bc[5] == 'N' ? uc : -uc);
        const double uxm = gxm * uc_p[o + oxm], uxp = gxp * uc_p[o + oxp];
        const double uym = gym * uc_p[o + oym], uyp = gyp * uc_p[o + oyp];
        const double Lu = (uxp - 2.0 * uc + uxm) * a.Cx + (uyp - 2.0 * uc + uym) * a.Cy + (up - 2.0 * uc + um);
        // effective right-hand side: rhs * dz^2 minus the Dirichlet boundary term
        double v = f_p[o];
        v = v * a.c3;
        v = v * a.c3;
        if (cxy != 0.0) v = v - cxy * ez[z];
        if (z == 0) v = v - czlo;
        if (z == a.nz - 1) v = v - czhi;
        double res = v - Lu;
        if (z == 0) res += defect;
        r_p[o] = (TOut)res;
        um = uc;
        uc = up;
    };
    // a plain loop keeps the kernel at ~52 VGPRs (8 waves per SIMD), which hides the load
    // latency better than unrolling the march
    const int zend = min(z0 + RES_ZC, a.nz);
    for (int z = z0; z < zend; z++) plane(z);
}

hipError_t launch_zfix64_f32(const ZArgs &a, const int4 *cols, int ncols, int nz_max, hipStream_t s)
{
    if (ncols == 0) return hipSuccess;
    const size_t lds = 2 * (size_t)nz_max * sizeof(double);
    hipLaunchKernelGGL(k_zfix64<float>, dim3(ncols), dim3(256), lds, s, a, cols);
    return hipGetLastError();
}

hipError_t launch_mode00(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                         double *plane_add, const double *upart, int nparts, int correct, hipStream_t s)
{
    const size_t lds = 2 * (size_t)nz * sizeof(double);
    hipLaunchKernelGGL(k_mode00, dim3(1), dim3(1024), lds, s, line_sums, ny, nz, b0, bn, inv_scale, plane_add, upart,
                       nparts, correct);
    return hipGetLastError();
}

hipError_t launch_mode00_batch(const Mode00Item *items, int nitems, int nz_max, int correct, hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    const size_t lds = 2 * (size_t)nz_max * sizeof(double);
    hipLaunchKernelGGL(k_mode00_batch, dim3(nitems), dim3(1024), lds, s, items, correct);
    return hipGetLastError();
}

int residual_parts(int nx, int ny) { return ((nx + 63) / 64) * ((ny + 3) / 4); }

int residual_zplanes() { return RES_ZC; }

int residual_blocks(int nx, int ny, int nz) { return residual_parts(nx, ny) * ((nz + RES_ZC - 1) / RES_ZC); }

hipError_t launch_residual_f32(const ResArgs &a, hipStream_t s)
{
    dim3 grid((unsigned)((a.nx + 63) / 64), (unsigned)((a.ny + 3) / 4), (unsigned)((a.nz + RES_ZC - 1) / RES_ZC));
    hipLaunchKernelGGL(k_residual<float>, grid, dim3(256), 0, s, a, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_residual_f32_batch(const ResArgs *solves, const int4 *items, int nitems, hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    ResArgs a{};
    hipLaunchKernelGGL(k_residual<float>, dim3((unsigned)nitems), dim3(256), 0, s, a, solves, items);
    return hipGetLastError();
}

}  // namespace pp
