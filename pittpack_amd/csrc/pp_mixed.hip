// pp_mixed.hip -- kernels of the mixed-precision form: FP32 transforms and z solves, with
// FP64 where FP32 cannot deliver FP64 accuracy after refinement.
//
//  * k_zfix64    : FP64 solve of the near-singular (kx, ky) columns (|eig+2| < tau) one
//                  column per block; the FP32 z solve leaves those columns untouched.  The
//                  system is thomasLowMem's without the pinning branch (src/triDiag.cpp:
//                  1693-1822), solved by parallel cyclic reduction in LDS.
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

#include <type_traits>

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
    // the column's rows in LDS: a, b, c, d (nz each)
    double *ra = zf, *rb = zf + a.nz, *rc = zf + 2 * a.nz, *rd = zf + 3 * a.nz;
    // element i of the column: in place (stride SX*SY) or in its z tile (stride W)
    long zs = (long)a.SX * a.SY;
    long off = (long)sc.y * a.SX + sc.x;
    if (a.tiled) {
        const int W = a.tile_w;
        const int rows = a.tile_rows ? a.tile_rows : (a.nz + 15) / 16 * 16;
        off = (((long)(sc.x / W) * a.SY + sc.y) * rows) * W + sc.x % W;
        zs = W;
    }
    R *data = reinterpret_cast<R *>(a.data);
    const int n = a.nz;
    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sc.x], a.Cx)), __dmul_rn(a.lamy[sc.y], a.Cy));
    const double b0 = a.zlo_dirichlet ? eig - 1. : eig + 1.;
    const double bn = a.zhi_dirichlet ? eig - 1. : eig + 1.;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const Row3d r = zrow(i, n, eig, b0, bn);
        ra[i] = r.a;
        rb[i] = r.b;
        rc[i] = r.c;
        rd[i] = (double)data[off + (long)i * zs];
    }
    __syncthreads();
    // parallel cyclic reduction: step S eliminates rows i - S and i + S from row i (rows past
    // either end are identity rows); after the step with S >= n / 2 every row stands alone
    constexpr int RPT = 4;  // rows per thread (n <= 4 * 256)
    for (int S = 1; S < n; S <<= 1) {
        double na[RPT], nb[RPT], nc[RPT], nd[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const int i = threadIdx.x + k * (int)blockDim.x;
            if (i < n) {
                const bool lo = i - S >= 0, hi = i + S < n;
                const double k1 = lo ? ra[i] / rb[i - S] : 0.0;
                const double k2 = hi ? rc[i] / rb[i + S] : 0.0;
                na[k] = lo ? -ra[i - S] * k1 : 0.0;
                nc[k] = hi ? -rc[i + S] * k2 : 0.0;
                nb[k] = rb[i] - (lo ? rc[i - S] * k1 : 0.0) - (hi ? ra[i + S] * k2 : 0.0);
                nd[k] = rd[i] - (lo ? rd[i - S] * k1 : 0.0) - (hi ? rd[i + S] * k2 : 0.0);
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const int i = threadIdx.x + k * (int)blockDim.x;
            if (i < n) {
                ra[i] = na[k];
                rb[i] = nb[k];
                rc[i] = nc[k];
                rd[i] = nd[k];
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) data[off + (long)i * zs] = (R)(rd[i] / rb[i]);
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
// Periodic z (zper): the column is the reference's singular cyclic one
// (shermanMorrisonThomasV1, src/triDiag.cpp:2112-2255), i.e. the cyclic operator with row 0
// replaced by u0 + u_{N-1} = (r0 + r_{N-1})/2, solved by the same Sherman-Morrison sweeps as
// k_smt_v1; plane_add[nz] = (u_{N-1} - 2 u0 + u1 - S0)/(NxNy) and plane_add[nz+1] = the
// row-0 target (S0 + S_{N-1})/2; in a correction, row 0 is that target minus the sum of the
// current u over planes 0 and N-1 (the residual kernel's partial sums take both planes).
__device__ void mode00_body(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                            double *plane_add, const double *upart, int nparts, int correct, int zper)
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
    if (threadIdx.x == 0 && zper) {
        const double s0 = r[0];
        const double target = correct ? plane_add[nz + 1] : (r[0] + r[nz - 1]) * 0.5;
        r[0] = correct ? target - red[0] : target;
        // thomasLowMemNoBCV1 sweeps of shermanMorrisonThomasV1 (b = {1, -2, -2}, gamma = -1,
        // alpha = beta = 1): row 0 decoupled with diagonal 2, interior 1 -2 1, last row -1
        const double bb[3] = {2.0, -2.0, -1.0};
        auto sweep = [&](double *x) {
            double bet = bb[0];
            x[0] = x[0] / bet;
            cp[1] = 0.0;
            bet = bb[1];
            x[1] = (x[1] - x[0]) / bet;
            for (int j = 2; j < nz - 1; j++) {
                cp[j] = 1.0 / bet;
                bet = bb[1] - cp[j];
                x[j] = (x[j] - x[j - 1]) / bet;
            }
            cp[nz - 1] = 1.0 / bet;
            bet = bb[2] - cp[nz - 1];
            x[nz - 1] = (x[nz - 1] - x[nz - 2]) / bet;
            for (int j = nz - 2; j >= 0; j--) x[j] -= cp[j + 1] * x[j + 1];
        };
        double *q = red;  // the Sherman-Morrison vector (the reduction is done with red)
        sweep(r);
        for (int i = 0; i < nz; i++) q[i] = 0.0;
        q[0] = -1.0;
        q[nz - 1] = 1.0;
        sweep(q);
        const double fact = (r[0] - r[nz - 1]) / (1. + q[0] - q[nz - 1]);
        for (int i = 0; i < nz; i++) r[i] -= fact * q[i];
        if (!correct) {
            plane_add[nz] = (r[nz - 1] - 2.0 * r[0] + r[1] - s0) * inv_scale;
            plane_add[nz + 1] = target;
        }
    } else if (threadIdx.x == 0) {
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
                                                 int correct, int zper)
{
    mode00_body(line_sums, ny, nz, b0, bn, inv_scale, plane_add, upart, nparts, correct, zper);
}

// batched: block b handles pinned member items[b]
__global__ void __launch_bounds__(1024) k_mode00_batch(const Mode00Item *items, int correct)
{
    const Mode00Item it = items[blockIdx.x];
    mode00_body(it.line_sums, it.ny, it.nz, it.b0, it.bn, it.inv_scale, it.plane_add, it.upart, it.nparts, correct,
                it.zper);
}

constexpr int RES_ZC = 30;  // planes marched per thread (a multiple of the 3-plane unroll)
constexpr int RES_TX = 64;  // block tile: one 64-lane wave per row ...
constexpr int RES_TY = 4;   // ... four rows

// Residual of one block tile (64 columns x 4 rows), marching RES_ZC planes with u(z-1), u(z),
// u(z+1) in registers.  Each wave reads one 512-byte row per field and plane (the access
// shape that streams fastest here); the four rows of u(z) are shared through a
// double-buffered LDS tile, so in-plane neighbours inside the tile are LDS reads and only the
// tile's edge lanes (x) and edge waves (y) need a halo or ghost value from memory: four
// loads per cell -- u(z+1), f and one halo per direction -- instead of seven.
// Batched launches: block b works on items[b] = {member, bx, by, bz} with solves[member].
// TWO: grids with nx = 1 (mod 64) or ny = 1 (mod 4), whose last column (row) is the first lane
// (wave) of its tile and so needs both neighbours of that direction from memory: one more
// halo load per direction and plane.
template <typename TOut, bool TWO, int SRC>
__global__ void __launch_bounds__(256) k_residual(ResArgs a0, const ResArgs *solves, const int4 *items)
{
    ResArgs a = a0;
    // Blocks are dealt to the 8 XCDs round-robin and each XCD has its own L2: numbered so
    // that each XCD works on a contiguous run of tiles, the halo lines a tile shares with its
    // neighbours are fetched once per XCD instead of once per tile
    const int lin = xcd_logical(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                gridDim.x * gridDim.y * gridDim.z);
    int bx = lin % gridDim.x, by = (lin / gridDim.x) % gridDim.y, bz = lin / (gridDim.x * gridDim.y);
    int gx = gridDim.x;
    if (items) {
        const int4 it = items[lin];
        a = solves[it.x];
        bx = it.y;
        by = it.z;
        bz = it.w;
        gx = (a.nx + RES_TX - 1) / RES_TX;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = bx * RES_TX + lane;
    const int y = by * RES_TY + w;
    const int z0 = bz * RES_ZC;
    const bool valid = x < a.nx && y < a.ny;
    const long sy = a.nx, sz = (long)a.nx * a.ny;
    // global address space: generic (flat) accesses force a full wait on every load
    using gdouble = const __attribute__((address_space(1))) double;
    using gout = __attribute__((address_space(1))) TOut;
    // lanes outside the grid work on the clamped cell (in bounds) and store nothing
    const int xc = min(x, a.nx - 1), yc = min(y, a.ny - 1);
    const long cell = xc + sy * yc;
    // the current solution: FP64, or its FP32 parts summed in FP64 (SRC, ResSource)
    using gfloat = const __attribute__((address_space(1))) float;
    gdouble *uc_p = SRC == RES_U64 ? (gdouble *)(a.u + cell) : nullptr;
    gfloat *u0_p = SRC != RES_U64 ? (gfloat *)(a.u0 + cell) : nullptr;
    gfloat *u1_p = SRC == RES_U0_U1 ? (gfloat *)(a.u1 + cell) : nullptr;
    auto ldu = [&](long o) -> double {
        if constexpr (SRC == RES_U64) return uc_p[o];
        else if constexpr (SRC == RES_U0) return (double)u0_p[o];
        else return (double)u0_p[o] + (double)u1_p[o];
    };
    gdouble *f_p = (gdouble *)(a.f + cell);
    __shared__ double red[256];
    if (a.upart && bz == 0) {
        // plane-0 sum of u (periodic z: planes 0 and nz-1), one partial per (x, y) tile
        red[threadIdx.x] = valid ? ldu(0) + (a.bc[4] == 'P' ? ldu(sz * (a.nz - 1)) : 0.0) : 0.0;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.upart[by * gx + bx] = red[0];
    }
    const double *ex = a.dtab, *ey = a.dtab + a.nx, *ez = a.dtab + a.nx + a.ny;
    const double *ef = a.dtab + a.nx + a.ny + a.nz;
    // Halo lanes and waves: the neighbour is outside the tile, or outside the grid where the
    // ghost closure the spectral solve diagonalises applies (P wraps, N mirrors, D is
    // antisymmetric).  Only a cell that is both the first lane (wave) of its tile and the
    // grid's last column (row) needs both neighbours of one direction from memory (TWO);
    // otherwise one halo load per direction serves every lane (a lane with no halo in a
    // direction reads its own cell): unconditional loads keep the compiler's count of
    // outstanding loads exact, which the pipelined march depends on.
    auto ghost = [](char bc, long wrap, long &off, double &sg) {
        off = bc == 'P' ? wrap : 0;
        sg = bc == 'D' ? -1.0 : 1.0;
    };
    const bool hxm = lane == 0, hxp = lane == 63 || xc == a.nx - 1;
    const bool hym = w == 0, hyp = w == RES_TY - 1 || yc == a.ny - 1;
    long oxm = -1, oxp = 1, oym = -sy, oyp = sy;
    double gxm = 1.0, gxp = 1.0, gym = 1.0, gyp = 1.0;
    if (xc == 0) ghost(a.bc[0], a.nx - 1, oxm, gxm);
    if (xc == a.nx - 1) ghost(a.bc[1], -(long)(a.nx - 1), oxp, gxp);
    if (yc == 0) ghost(a.bc[2], sy * (a.ny - 1), oym, gym);
    if (yc == a.ny - 1) ghost(a.bc[3], -sy * (a.ny - 1), oyp, gyp);
    const long ohx = hxm ? oxm : (hxp ? oxp : 0), ohy = hym ? oym : (hyp ? oyp : 0);
    // the upper neighbour of a cell whose lower one already takes the halo load
    const bool bothx = TWO && hxm && hxp, bothy = TWO && hym && hyp;
    const long ohx2 = bothx ? oxp : 0, ohy2 = bothy ? oyp : 0;
    // z-profile of the x/y-face Dirichlet terms for this block's planes, and the u(z) tile
    __shared__ double ezs[RES_ZC];
    __shared__ double tile[2][RES_TY][RES_TX];
    if (threadIdx.x < RES_ZC) ezs[threadIdx.x] = z0 + (int)threadIdx.x < a.nz ? gld(ez, z0 + threadIdx.x) : 0.0;
    // Dirichlet terms of the effective right-hand side: x and y faces are a per-(x, y)
    // constant times ez[z], z faces a constant on the first and last plane
    double cxy = 0.0, czlo = 0.0, czhi = 0.0;
    if (a.dmask) {
        if ((a.dmask & 1) && xc == 0) cxy += 2. * ef[0] * ey[yc];
        if ((a.dmask & 2) && xc == a.nx - 1) cxy += 2. * ef[1] * ey[yc];
        if ((a.dmask & 4) && yc == 0) cxy += 2. * ex[xc] * ef[2];
        if ((a.dmask & 8) && yc == a.ny - 1) cxy += 2. * ex[xc] * ef[3];
        if (a.dmask & 16) czlo = 2. * ex[xc] * ey[yc] * ef[4];
        if (a.dmask & 32) czhi = 2 * ex[xc] * ey[yc] * ef[5];
    }
    const double defect = a.pin_defect ? a.pin_defect[0] : 0.0;
    gout *r_p = (gout *)(reinterpret_cast<TOut *>(a.r) + cell);

    double uc = ldu(sz * z0);
    double um;
    if (z0 > 0) um = ldu(sz * (z0 - 1));
    else um = a.bc[4] == 'P' ? ldu(sz * (a.nz - 1)) : (a.bc[4] == 'N' ? uc : -uc);
    // loads of one plane: u(z+1) (at the top face: plane 0 for P, else this plane, replaced
    // by the ghost below), the right-hand side and the halo value per direction
    struct Plane {
        double up, f, hx, hy, hx2, hy2;
    };
    const bool ztop_wrap = a.bc[5] == 'P';
    auto fetch = [&](int z, Plane &q) {
        const long o = sz * z;
        q.up = ldu(z + 1 < a.nz ? o + sz : (ztop_wrap ? 0 : o));
        q.f = f_p[o];
        q.hx = ldu(o + ohx);
        q.hy = ldu(o + ohy);
        if constexpr (TWO) {
            q.hx2 = ldu(o + ohx2);
            q.hy2 = ldu(o + ohy2);
        } else {
            q.hx2 = q.hx;
            q.hy2 = q.hy;
        }
    };
    // software pipelined march: the loads of planes z + 1 and z + 2 are in flight while plane
    // z is computed and stored, so two planes of latency overlap this plane's arithmetic.  The
    // loop body has no branch around a memory access (neighbours are read from LDS
    // unconditionally and selected; a tile wholly inside the grid stores unconditionally), and
    // a block of RES_ZC planes runs unrolled by three over three plane buffers, so no loaded
    // register is copied (a copy would wait for its load) and the compiler can count the loads
    // in flight instead of waiting for all of them.
    const int zend = min(z0 + RES_ZC, a.nz);
    const bool full_tile = (bx + 1) * RES_TX <= a.nx && (by + 1) * RES_TY <= a.ny;
    auto step = [&](int z, const Plane &cur, auto all_valid) {
        constexpr bool ALLV = decltype(all_valid)::value;
        const int b = (z - z0) & 1;
        tile[b][w][lane] = uc;
        __syncthreads();  // one barrier per plane: the other buffer was last read before it
        const long o = sz * z;
        const double t_xm = tile[b][w][lane > 0 ? lane - 1 : 0];
        const double t_xp = tile[b][w][lane < 63 ? lane + 1 : 63];
        const double t_ym = tile[b][w > 0 ? w - 1 : 0][lane];
        const double t_yp = tile[b][w < RES_TY - 1 ? w + 1 : w][lane];
        const double up = (z + 1 < a.nz || ztop_wrap) ? cur.up : (a.bc[5] == 'N' ? uc : -uc);
        const double uxm = hxm ? gxm * cur.hx : t_xm;
        const double uxp = hxp ? gxp * (bothx ? cur.hx2 : cur.hx) : t_xp;
        const double uym = hym ? gym * cur.hy : t_ym;
        const double uyp = hyp ? gyp * (bothy ? cur.hy2 : cur.hy) : t_yp;
        // second differences as differences of first differences: neighbouring values are
        // close, so each first difference is exact (Sterbenz) and the rounding scales with the
        // differences, not with u
        const double Lu = ((uxp - uc) - (uc - uxm)) * a.Cx + ((uyp - uc) - (uc - uym)) * a.Cy + ((up - uc) - (uc - um));
        // effective right-hand side: rhs * dz^2 minus the Dirichlet boundary term
        double v = cur.f;
        v = v * a.c3;
        v = v * a.c3;
        v = v - cxy * ezs[z - z0];  // cxy = 0 away from x / y Dirichlet faces
        if (z == 0) v = v - czlo;
        if (z == a.nz - 1) v = v - czhi;
        double res = v - Lu;
        if (z == 0) res += defect;
        if (ALLV || valid) r_p[o] = (TOut)res;
        um = uc;
        uc = up;
    };
    if (full_tile && zend - z0 == RES_ZC) {
        Plane A, B, C;
        fetch(z0, A);
        fetch(z0 + 1, B);
        // fully unrolled: no loop back edge, so the counts of loads in flight stay exact
#pragma clang loop unroll(full)
        for (int g = 0; g < RES_ZC / 3; g++) {
            const int z = z0 + 3 * g;
            fetch(min(z + 2, zend - 1), C);
            step(z, A, std::true_type{});
            fetch(min(z + 3, zend - 1), A);
            step(z + 1, B, std::true_type{});
            fetch(min(z + 4, zend - 1), B);
            step(z + 2, C, std::true_type{});
        }
    } else {
        // edge tiles and the last, shorter block of planes
        Plane cur;
        fetch(z0, cur);
        for (int z = z0; z < zend; z++) {
            Plane nxt;
            fetch(min(z + 1, zend - 1), nxt);
            step(z, cur, std::false_type{});
            cur = nxt;
        }
    }
}

// Residual with two adjacent x cells per lane (grids with even nx): a block tile is 128
// columns x 4 rows, u and f are read as one 16-byte load per lane (one 1 KiB row per wave and
// field), r is stored as one 8-byte pair, and the u(z) tile in LDS holds pairs -- half the
// load and LDS instructions per cell of k_residual, whose address processing bounds it.
// Each lane takes its left and right halo value from memory (a lane with no halo in a
// direction reads its own cell); the rest is k_residual's march.  Plane-0 partial sums of u
// are written per 64-column half tile, in k_residual's layout.  TWOY: ny = 1 (mod 4).
typedef double rp_d2 __attribute__((ext_vector_type(2)));
typedef float rp_f2 __attribute__((ext_vector_type(2)));
constexpr int RP_CX = 2 * RES_TX;  // columns per block tile

template <bool TWOY, int SRC>
__global__ void __launch_bounds__(256) k_residual_pair(ResArgs a0, const ResArgs *solves, const int4 *items)
{
    ResArgs a = a0;
    const int lin = xcd_logical(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                gridDim.x * gridDim.y * gridDim.z);
    int bx = lin % gridDim.x, by = (lin / gridDim.x) % gridDim.y, bz = lin / (gridDim.x * gridDim.y);
    if (items) {
        // batched: block b works on items[b] = {member, bx, by, bz} (128-column tiles)
        const int4 it = items[lin];
        a = solves[it.x];
        bx = it.y;
        by = it.z;
        bz = it.w;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = bx * RP_CX + 2 * lane;  // first cell of the lane's pair
    const int y = by * RES_TY + w;
    const int z0 = bz * RES_ZC;
    const bool valid = x < a.nx && y < a.ny;
    const long sy = a.nx, sz = (long)a.nx * a.ny;
    using gdouble = const __attribute__((address_space(1))) double;
    using gd2 = const __attribute__((address_space(1))) rp_d2;
    using gf2 = __attribute__((address_space(1))) rp_f2;
    // lanes outside the grid work on the clamped pair (in bounds) and store nothing; nx is
    // even, so a pair is wholly inside or wholly outside and every pair starts 16-byte aligned
    const int xc = min(x, a.nx - 2), yc = min(y, a.ny - 1);
    const long cell = xc + sy * yc;
    // every offset below counts doubles (floats for r) from the pair's first cell; pairs are
    // read and written through a vector cast of the element address
    // the current solution: FP64, or its FP32 parts summed in FP64 (SRC, ResSource)
    using gfloat = const __attribute__((address_space(1))) float;
    using gpf2 = const __attribute__((address_space(1))) rp_f2;
    const double *u_c = SRC == RES_U64 ? a.u + cell : nullptr, *f_c = a.f + cell;
    const float *u0_c = SRC != RES_U64 ? a.u0 + cell : nullptr;
    const float *u1_c = SRC == RES_U0_U1 ? a.u1 + cell : nullptr;
    auto ldu = [&](long o) -> double {
        if constexpr (SRC == RES_U64) return *(gdouble *)(u_c + o);
        else if constexpr (SRC == RES_U0) return (double)*(gfloat *)(u0_c + o);
        else return (double)*(gfloat *)(u0_c + o) + (double)*(gfloat *)(u1_c + o);
    };
    auto ldu2 = [&](long o) -> rp_d2 {
        if constexpr (SRC == RES_U64) {
            return *(gd2 *)(u_c + o);
        } else {
            const rp_f2 p = *(gpf2 *)(u0_c + o);
            rp_d2 v = {(double)p.x, (double)p.y};
            if constexpr (SRC == RES_U0_U1) {
                const rp_f2 q = *(gpf2 *)(u1_c + o);
                v.x += (double)q.x;
                v.y += (double)q.y;
            }
            return v;
        }
    };
    auto ldf2 = [&](long o) { return *(gd2 *)(f_c + o); };
    __shared__ double red[256];
    if (a.upart && bz == 0) {
        // plane-0 (periodic z: and plane nz-1) sums of u per 64-column half tile
        double sum = 0.0;
        if (valid) {
            const rp_d2 v = ldu2(0);
            sum = v.x + v.y;
            if (a.bc[4] == 'P') {
                const rp_d2 q = ldu2(sz * (a.nz - 1));
                sum += q.x + q.y;
            }
        }
        red[(lane >> 5) * 128 + w * 32 + (lane & 31)] = sum;
        __syncthreads();
        for (int st = 64; st > 0; st >>= 1) {
            if ((int)(threadIdx.x & 127) < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        const int g64 = (a.nx + RES_TX - 1) / RES_TX;
        if (threadIdx.x == 0) a.upart[by * g64 + 2 * bx] = red[0];
        if (threadIdx.x == 128 && 2 * bx + 1 < g64) a.upart[by * g64 + 2 * bx + 1] = red[128];
    }
    const double *ex = a.dtab, *ey = a.dtab + a.nx, *ez = a.dtab + a.nx + a.ny;
    const double *ef = a.dtab + a.nx + a.ny + a.nz;
    // halo offsets from the pair's first cell: left neighbour of cell xc, right neighbour of
    // cell xc + 1; at the grid's faces the ghost closure (P wraps, N mirrors, D antisymmetric)
    auto ghost = [](char bc, long wrap, long self, long &off, double &sg) {
        off = bc == 'P' ? wrap : self;
        sg = bc == 'D' ? -1.0 : 1.0;
    };
    const bool hxm = lane == 0, hxp = lane == RES_TX - 1 || xc + 1 == a.nx - 1;
    const bool hym = w == 0, hyp = w == RES_TY - 1 || yc == a.ny - 1;
    long oxm = -1, oxp = 2, oym = -sy, oyp = sy;
    double gxm = 1.0, gxp = 1.0, gym = 1.0, gyp = 1.0;
    if (xc == 0) ghost(a.bc[0], a.nx - 1, 0, oxm, gxm);
    if (xc + 1 == a.nx - 1) ghost(a.bc[1], -(long)(a.nx - 2), 1, oxp, gxp);
    if (yc == 0) ghost(a.bc[2], sy * (a.ny - 1), 0, oym, gym);
    if (yc == a.ny - 1) ghost(a.bc[3], -sy * (a.ny - 1), 0, oyp, gyp);
    const long ohl = hxm ? oxm : 0, ohr = hxp ? oxp : 0;
    const long ohy = hym ? oym : (hyp ? oyp : 0);
    const bool bothy = TWOY && hym && hyp;
    const long ohy2 = bothy ? oyp : 0;
    __shared__ double ezs[RES_ZC];
    __shared__ rp_d2 tile[2][RES_TY][RES_TX];
    if (threadIdx.x < RES_ZC) ezs[threadIdx.x] = z0 + (int)threadIdx.x < a.nz ? gld(ez, z0 + threadIdx.x) : 0.0;
    // Dirichlet terms per cell: x and y faces a per-(x, y) constant times ez[z], z faces a
    // constant on the first and last plane
    rp_d2 cxy = {0.0, 0.0}, czlo = {0.0, 0.0}, czhi = {0.0, 0.0};
    if (a.dmask) {
        const double ex0 = ex[xc], ex1 = ex[xc + 1], eyc = ey[yc];
        if ((a.dmask & 1) && xc == 0) cxy.x += 2. * ef[0] * eyc;
        if ((a.dmask & 2) && xc + 1 == a.nx - 1) cxy.y += 2. * ef[1] * eyc;
        if ((a.dmask & 4) && yc == 0) {
            cxy.x += 2. * ex0 * ef[2];
            cxy.y += 2. * ex1 * ef[2];
        }
        if ((a.dmask & 8) && yc == a.ny - 1) {
            cxy.x += 2. * ex0 * ef[3];
            cxy.y += 2. * ex1 * ef[3];
        }
        if (a.dmask & 16) {
            czlo.x = 2. * ex0 * eyc * ef[4];
            czlo.y = 2. * ex1 * eyc * ef[4];
        }
        if (a.dmask & 32) {
            czhi.x = 2 * ex0 * eyc * ef[5];
            czhi.y = 2 * ex1 * eyc * ef[5];
        }
    }
    const double defect = a.pin_defect ? a.pin_defect[0] : 0.0;
    float *r_c = reinterpret_cast<float *>(a.r) + cell;

    rp_d2 uc = ldu2(sz * z0);
    rp_d2 um;
    if (z0 > 0) um = ldu2(sz * (z0 - 1));
    else um = a.bc[4] == 'P' ? ldu2(sz * (a.nz - 1)) : (a.bc[4] == 'N' ? uc : -uc);
    struct Plane {
        rp_d2 up, f, hy, hy2;
        double hl, hr;
    };
    const bool ztop_wrap = a.bc[5] == 'P';
    auto fetch = [&](int z, Plane &q) {
        const long o = sz * z;
        q.up = ldu2(z + 1 < a.nz ? o + sz : (ztop_wrap ? 0 : o));
        q.f = ldf2(o);
        q.hl = ldu(o + ohl);
        q.hr = ldu(o + ohr);
        q.hy = ldu2(o + ohy);
        if constexpr (TWOY) q.hy2 = ldu2(o + ohy2);
        else q.hy2 = q.hy;
    };
    const int zend = min(z0 + RES_ZC, a.nz);
    const bool full_tile = (bx + 1) * RP_CX <= a.nx && (by + 1) * RES_TY <= a.ny;
    auto step = [&](int z, const Plane &cur, auto all_valid) {
        constexpr bool ALLV = decltype(all_valid)::value;
        const int b = (z - z0) & 1;
        tile[b][w][lane] = uc;
        __syncthreads();  // one barrier per plane: the other buffer was last read before it
        const double t_l = tile[b][w][lane > 0 ? lane - 1 : 0].y;
        const double t_r = tile[b][w][lane < RES_TX - 1 ? lane + 1 : RES_TX - 1].x;
        const rp_d2 t_ym = tile[b][w > 0 ? w - 1 : 0][lane];
        const rp_d2 t_yp = tile[b][w < RES_TY - 1 ? w + 1 : w][lane];
        const rp_d2 up = (z + 1 < a.nz || ztop_wrap) ? cur.up : (a.bc[5] == 'N' ? uc : -uc);
        const double uxl = hxm ? gxm * cur.hl : t_l;
        const double uxr = hxp ? gxp * cur.hr : t_r;
        const rp_d2 uym = hym ? gym * cur.hy : t_ym;
        const rp_d2 uyp = hyp ? gyp * (bothy ? cur.hy2 : cur.hy) : t_yp;
        // second differences as differences of first differences (as k_residual)
        const double Lu0 = ((uc.y - uc.x) - (uc.x - uxl)) * a.Cx + ((uyp.x - uc.x) - (uc.x - uym.x)) * a.Cy +
                           ((up.x - uc.x) - (uc.x - um.x));
        const double Lu1 = ((uxr - uc.y) - (uc.y - uc.x)) * a.Cx + ((uyp.y - uc.y) - (uc.y - uym.y)) * a.Cy +
                           ((up.y - uc.y) - (uc.y - um.y));
        rp_d2 v = cur.f;
        v = v * a.c3;
        v = v * a.c3;
        v = v - cxy * ezs[z - z0];
        if (z == 0) v = v - czlo;
        if (z == a.nz - 1) v = v - czhi;
        double res0 = v.x - Lu0, res1 = v.y - Lu1;
        if (z == 0) {
            res0 += defect;
            res1 += defect;
        }
        if (ALLV || valid) *(gf2 *)(r_c + sz * z) = rp_f2{(float)res0, (float)res1};
        um = uc;
        uc = up;
    };
    if (full_tile && zend - z0 == RES_ZC) {
        Plane A, B, C;
        fetch(z0, A);
        fetch(z0 + 1, B);
#pragma clang loop unroll(full)
        for (int g = 0; g < RES_ZC / 3; g++) {
            const int z = z0 + 3 * g;
            fetch(min(z + 2, zend - 1), C);
            step(z, A, std::true_type{});
            fetch(min(z + 3, zend - 1), A);
            step(z + 1, B, std::true_type{});
            fetch(min(z + 4, zend - 1), B);
            step(z + 2, C, std::true_type{});
        }
    } else {
        Plane cur;
        fetch(z0, cur);
        for (int z = z0; z < zend; z++) {
            Plane nxt;
            fetch(min(z + 1, zend - 1), nxt);
            step(z, cur, std::false_type{});
            cur = nxt;
        }
    }
}

hipError_t launch_zfix64_f32(const ZArgs &a, const int4 *cols, int ncols, int nz_max, hipStream_t s)
{
    if (ncols == 0) return hipSuccess;
    if (nz_max > 4 * 256) return hipErrorInvalidValue;  // four rows per thread
    const size_t lds = 4 * (size_t)nz_max * sizeof(double);
    hipLaunchKernelGGL(k_zfix64<float>, dim3(ncols), dim3(256), lds, s, a, cols);
    return hipGetLastError();
}

hipError_t launch_mode00(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                         double *plane_add, const double *upart, int nparts, int correct, int zper, hipStream_t s)
{
    if (nz > 1024 || (zper && nz < 3)) return hipErrorInvalidValue;  // red[] holds the periodic column's vector
    const size_t lds = 2 * (size_t)nz * sizeof(double);
    hipLaunchKernelGGL(k_mode00, dim3(1), dim3(1024), lds, s, line_sums, ny, nz, b0, bn, inv_scale, plane_add, upart,
                       nparts, correct, zper);
    return hipGetLastError();
}

hipError_t launch_mode00_batch(const Mode00Item *items, int nitems, int nz_max, int correct, hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    if (nz_max > 1024) return hipErrorInvalidValue;
    const size_t lds = 2 * (size_t)nz_max * sizeof(double);
    hipLaunchKernelGGL(k_mode00_batch, dim3(nitems), dim3(1024), lds, s, items, correct);
    return hipGetLastError();
}

// The pair kernel's 128-column tiles leave lanes idle on narrow grids (nx = 32: a quarter of
// them busy); it is taken where it keeps at least as many lanes busy as the 64-column kernel
bool residual_pairs(int nx)
{
    if (nx % 2 != 0 || nx < 2) return false;
    const long pair_cols = (long)(nx + RP_CX - 1) / RP_CX * RP_CX, cols = (long)(nx + RES_TX - 1) / RES_TX * RES_TX;
    return pair_cols <= cols;
}

int residual_tile_x_of(int nx) { return residual_pairs(nx) ? RP_CX : RES_TX; }

template <bool TWOY>
static hipError_t launch_pair_src(dim3 grid, const ResArgs &a, const ResArgs *solves, const int4 *items, int src,
                                  hipStream_t s)
{
    switch (src) {
    case RES_U0: hipLaunchKernelGGL((k_residual_pair<TWOY, RES_U0>), grid, dim3(256), 0, s, a, solves, items); break;
    case RES_U0_U1: hipLaunchKernelGGL((k_residual_pair<TWOY, RES_U0_U1>), grid, dim3(256), 0, s, a, solves, items); break;
    default: hipLaunchKernelGGL((k_residual_pair<TWOY, RES_U64>), grid, dim3(256), 0, s, a, solves, items); break;
    }
    return hipGetLastError();
}

template <bool TWO>
static hipError_t launch_res_src(dim3 grid, const ResArgs &a, const ResArgs *solves, const int4 *items, int src,
                                 hipStream_t s)
{
    switch (src) {
    case RES_U0: hipLaunchKernelGGL((k_residual<float, TWO, RES_U0>), grid, dim3(256), 0, s, a, solves, items); break;
    case RES_U0_U1:
        hipLaunchKernelGGL((k_residual<float, TWO, RES_U0_U1>), grid, dim3(256), 0, s, a, solves, items);
        break;
    default: hipLaunchKernelGGL((k_residual<float, TWO, RES_U64>), grid, dim3(256), 0, s, a, solves, items); break;
    }
    return hipGetLastError();
}

hipError_t launch_residual_pair_f32_batch(const ResArgs *solves, const int4 *items, int nitems, bool twoy, int src,
                                          hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    ResArgs a{};
    return twoy ? launch_pair_src<true>(dim3((unsigned)nitems), a, solves, items, src, s)
                : launch_pair_src<false>(dim3((unsigned)nitems), a, solves, items, src, s);
}

int residual_parts(int nx, int ny) { return ((nx + RES_TX - 1) / RES_TX) * ((ny + RES_TY - 1) / RES_TY); }

int residual_tile_x() { return RES_TX; }

int residual_tile_y() { return RES_TY; }

int residual_zplanes() { return RES_ZC; }

int residual_blocks(int nx, int ny, int nz) { return residual_parts(nx, ny) * ((nz + RES_ZC - 1) / RES_ZC); }

bool residual_two_halo(int nx, int ny) { return nx % RES_TX == 1 || ny % RES_TY == 1; }

hipError_t launch_residual_f32(const ResArgs &a, int src, hipStream_t s)
{
    if (residual_pairs(a.nx)) {
        const dim3 grid((unsigned)((a.nx + RP_CX - 1) / RP_CX), (unsigned)((a.ny + RES_TY - 1) / RES_TY),
                        (unsigned)((a.nz + RES_ZC - 1) / RES_ZC));
        return a.ny % RES_TY == 1 ? launch_pair_src<true>(grid, a, nullptr, nullptr, src, s)
                                  : launch_pair_src<false>(grid, a, nullptr, nullptr, src, s);
    }
    dim3 grid((unsigned)((a.nx + RES_TX - 1) / RES_TX), (unsigned)((a.ny + RES_TY - 1) / RES_TY),
              (unsigned)((a.nz + RES_ZC - 1) / RES_ZC));
    return residual_two_halo(a.nx, a.ny) ? launch_res_src<true>(grid, a, nullptr, nullptr, src, s)
                                         : launch_res_src<false>(grid, a, nullptr, nullptr, src, s);
}

hipError_t launch_residual_f32_batch(const ResArgs *solves, const int4 *items, int nitems, bool two, int src,
                                     hipStream_t s)
{
    if (nitems == 0) return hipSuccess;
    ResArgs a{};
    return two ? launch_res_src<true>(dim3((unsigned)nitems), a, solves, items, src, s)
               : launch_res_src<false>(dim3((unsigned)nitems), a, solves, items, src, s);
}

}  // namespace pp
