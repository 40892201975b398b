// pp_zsolve.h -- batched tridiagonal z solve for CDNA4 (gfx950).
//
// k_zsolve: one (kx, ky) column per 16-lane DPP row; PCR across the 16 lanes (DPP row
// rotations, no LDS) splits the column into 16 interleaved systems, then each lane runs
// Thomas on its own rows.  Replaces solveThmBatch + fillInArrayContig/Back + thomasLowMem
// (src/pencilDcmp.cpp:5062-5234, src/triDiag.cpp:1629-1822).
//
// The spectral field keeps the size of the physical field (half-complex layout on periodic
// axes), so column (sx, sy) uses the reference eigenvalue with j = sx, i = sy
// (getEigenVal, src/pencilDcmp.cpp:4393-4417).
#pragma once
#include <hip/hip_runtime.h>

#include "pp_xcd.h"

namespace pp {

// ---------------------------------------------------------------------------------------
// z solve
// ---------------------------------------------------------------------------------------
struct ZItem;
struct ZSolve;

struct ZArgs {
    void *data;  // [nz][SY][SX], in place
    int SX, SY, nz;
    const double *lamx;  // -2 + 2 cos(...) per x mode (SX)
    const double *lamy;  // per y mode (SY)
    double Cx, Cy;
    int zlo_dirichlet, zhi_dirichlet;  // 1 = D, 0 = N
    int zero_mode;                      // 0: reference row-0 rule; 1: only when singular (z NN)
    // mixed precision: columns with |eig+2| < tau (and not pinned) are left untouched for the
    // FP64 column solve; pinned_zero: the pinned (0,0) column is written as zeros (its
    // contribution is solved exactly in FP64 elsewhere)
    double tau;
    int pinned_zero;
    // batched launch (null: one solve): block b solves the columns of items[b]
    const ZItem *items;
    const ZSolve *solves;
    int nitems;
    // z-tiled layout (written by the y forward transform): tile (x/W, y) holds 16 J rows
    // (rows from nz on are zero padding) of W columns contiguously, element (x, y, z) at
    // ((x/W * SY + y) * 16 J + z) * W + x%W; W = tile_w (0: the block's CPB), a multiple of
    // CPB -- a block then takes CPB columns of a tile
    int tiled;
    int tile_w;
    int tile_rows;  // rows per tile (0: 16 J); more than 16 J pads each tile at its end
};

// one solve of a batched z launch: its field sits at `offset` in the launch's data arena
struct ZSolve {
    long offset;
    int SX, SY, nz;
    int zlo_dirichlet, zhi_dirichlet;
    double Cx, Cy;
    const double *lamx, *lamy;
};
struct ZItem {
    int solve, bx, by;
};

// Column element access.  The row stride zs is uniform, so row j of a lane sits at the
// uniform address data + j*zs (scalar registers) plus one per-lane 32-bit byte offset:
// this keeps one VGPR per lane for all rows instead of one 64-bit address per row.
// Valid while the field is below 4 GiB (the host picks the 64-bit form otherwise).
template <typename R, bool OFF32>
struct ColPtr {
    const char *base;
    long zsb;      // row stride in bytes
    uint32_t off;  // lane offset in bytes (OFF32)
    long off64;    // lane offset in bytes (!OFF32)
    __device__ __forceinline__ ColPtr(R *data, long zs, long lane_elem)
        : base(reinterpret_cast<const char *>(data)), zsb(zs * (long)sizeof(R)),
          off((uint32_t)(lane_elem * (long)sizeof(R))), off64(lane_elem * (long)sizeof(R)) {}
    __device__ __forceinline__ R *at(int j) const
    {
        const char *rowbase = base + (long)j * zsb;
        if constexpr (OFF32) return (R *)(rowbase + off);
        else return (R *)(rowbase + off64);
    }
};

// 1/x: hardware reciprocal refined by two Newton steps (within an ulp or two of 1/x; the
// pivots here are diagonally dominant, never near zero or denormal)
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float x)
{
    float r = __builtin_amdgcn_rcpf(x);
    float e = fmaf(-x, r, 1.0f);
    return fmaf(r, e, r);
}

// value of x from lane (t - S) mod 16 of this lane's 16-lane row
// (DPP row_ror:S; every source lane is valid, so no "old" operand is needed)
template <int S>
__device__ __forceinline__ float rot_from_lower(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x120 + S, 0xF, 0xF, true));
}
template <int S>
__device__ __forceinline__ double rot_from_lower(double x)
{
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_mov_dpp(lo, 0x120 + S, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, 0x120 + S, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// value of x from lane (t + S) mod 16
template <int S, typename R>
__device__ __forceinline__ R rot_from_upper(R x)
{
    return rot_from_lower<16 - S>(x);
}

template <typename R>
struct Row3 {
    R a, b, c;
};

template <int S, typename R>
__device__ __forceinline__ Row3<R> rotL3(Row3<R> r)
{
    return {rot_from_lower<S>(r.a), rot_from_lower<S>(r.b), rot_from_lower<S>(r.c)};
}
template <int S, typename R>
__device__ __forceinline__ Row3<R> rotR3(Row3<R> r)
{
    return {rot_from_upper<S>(r.a), rot_from_upper<S>(r.b), rot_from_upper<S>(r.c)};
}

// general PCR row update: new row from (self, left neighbour, right neighbour)
template <typename R>
__device__ __forceinline__ void pcr_row(Row3<R> &me, R &d, Row3<R> L, R dL, Row3<R> Rr, R dR)
{
    const R k1 = me.a * rcp_nr(L.b);
    const R k2 = me.c * rcp_nr(Rr.b);
    const R na = -L.a * k1;
    const R nc = -Rr.c * k2;
    const R nb = me.b - L.c * k1 - Rr.a * k2;
    d = d - dL * k1 - dR * k2;
    me = {na, nb, nc};
}

// One PCR step with stride S on a lane's J rows (row i = t + 16 j).  Rows j in
// {0, J-2, J-1} carry explicit coefficients (all rows when ALL); the others hold the
// column's Toeplitz coefficients (ta, tb, ta).
template <typename R, int J, bool ALL, int S>
__device__ __forceinline__ void pcr_step(R *d, Row3<R> *ex, R &ta, R &tb, int t)
{
    const bool hasL = t >= S;       // left neighbour in the same register row
    const bool hasR = t + S < 16;   // right neighbour in the same register row
    const Row3<R> Z0 = {(R)0, (R)1, (R)0};
    if constexpr (ALL) {
        // all rows explicit: fetch everything first (J <= 3, cheap)
        Row3<R> cl[J], cr[J];
        R dl[J], dr[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            cl[j] = rotL3<S>(ex[j]);
            cr[j] = rotR3<S>(ex[j]);
            dl[j] = rot_from_lower<S>(d[j]);
            dr[j] = rot_from_upper<S>(d[j]);
        }
#pragma unroll
        for (int j = 0; j < J; j++) {
            Row3<R> Lc = hasL ? cl[j] : (j > 0 ? cl[j - 1] : Z0);
            R Ld = hasL ? dl[j] : (j > 0 ? dl[j - 1] : (R)0);
            Row3<R> Rc = hasR ? cr[j] : (j + 1 < J ? cr[j + 1] : Z0);
            R Rd = hasR ? dr[j] : (j + 1 < J ? dr[j + 1] : (R)0);
            pcr_row(ex[j], d[j], Lc, Ld, Rc, Rd);
        }
    } else {
        // explicit rows: ex[0] = row j=0, ex[1] = j=J-2, ex[2] = j=J-1
        const Row3<R> Tt = {ta, tb, ta};
        const Row3<R> e0L = rotL3<S>(ex[0]), e0R = rotR3<S>(ex[0]);
        const Row3<R> eML = rotL3<S>(ex[1]), eMR = rotR3<S>(ex[1]);
        const Row3<R> eLL = rotL3<S>(ex[2]), eLR = rotR3<S>(ex[2]);
        const R kt = ta * rcp_nr(tb);
        R RLprev = (R)0;
        R RRcur = rot_from_upper<S>(d[0]);
#pragma unroll
        for (int j = 0; j < J; j++) {
            const R RLcur = rot_from_lower<S>(d[j]);
            const R RRnext = (j + 1 < J) ? rot_from_upper<S>(d[j + 1]) : (R)0;
            const R dL = hasL ? RLcur : (j > 0 ? RLprev : (R)0);
            const R dR = hasR ? RRcur : (j + 1 < J ? RRnext : (R)0);
            if (j == 0) {
                Row3<R> Lc = hasL ? e0L : Z0;
                Row3<R> Rc = hasR ? e0R : Tt;
                pcr_row(ex[0], d[0], Lc, dL, Rc, dR);
            } else if (j == J - 2) {
                Row3<R> Lc = hasL ? eML : Tt;
                Row3<R> Rc = hasR ? eMR : eLR;
                pcr_row(ex[1], d[j], Lc, dL, Rc, dR);
            } else if (j == J - 1) {
                Row3<R> Lc = hasL ? eLL : eML;
                Row3<R> Rc = hasR ? eLR : Z0;
                pcr_row(ex[2], d[j], Lc, dL, Rc, dR);
            } else {
                d[j] = d[j] - kt * (dL + dR);
            }
            RLprev = RLcur;
            RRcur = RRnext;
        }
        const R nta = -ta * kt;
        tb = tb - (R)2 * ta * kt;
        ta = nta;
    }
}

// ---------------------------------------------------------------------------------------
// LDS staging of a block's tile: 16 adjacent columns x all rows.  Global traffic moves in
// full 128-byte rows (16 columns x 8 B, four rows per wave instruction); the lanes then
// read their own rows from LDS.  Row stride 17 elements (odd) spreads a lane-strided
// column read over the banks.  `lrow(r)` maps tile row r to its LDS row.  CPB columns per
// block (16: 128-byte rows; 8: 64-byte rows, half the LDS per block).
// ---------------------------------------------------------------------------------------

template <typename R, int J, bool FULL, bool OFF32, int CPB, typename LRow, typename S = R>
__device__ __forceinline__ void ztile_load(R *tile, const ZArgs &a, int n, LRow lrow, int bx, int by)
{
    const int tid = threadIdx.x;
    const int c = tid % CPB, g = tid / CPB;
    if (a.tiled) {
        if (a.tile_w == 0 || a.tile_w == CPB) {
            // z tiles: 16 J rows (zero padding below nz) x CPB columns, contiguous; row r of
            // the tile is element r * CPB + c, so thread tid = g * CPB + c reads tid + q * 16 * CPB
            const S *__restrict__ src = reinterpret_cast<const S *>(a.data) + ((long)bx * a.SY + by) * (a.tile_rows ? a.tile_rows : 16 * J) * CPB;
#pragma unroll
            for (int q = 0; q < J; q++) tile[lrow(g + 16 * q) * (CPB + 1) + c] = (R)src[q * 16 * CPB + tid];
        } else {
            // wider tiles: this block's CPB columns of tile bx / (W / CPB), rows W apart
            const int W = a.tile_w, sub = W / CPB;
            const S *__restrict__ src = reinterpret_cast<const S *>(a.data) +
                                        ((long)(bx / sub) * a.SY + by) * (a.tile_rows ? a.tile_rows : 16 * J) * W + (bx % sub) * CPB + g * W + c;
#pragma unroll
            for (int q = 0; q < J; q++) tile[lrow(g + 16 * q) * (CPB + 1) + c] = (R)src[(long)q * 16 * W];
        }
        return;
    }
    const int col0 = bx * CPB;
    const bool cvalid = col0 + c < a.SX;
    const long zs = (long)a.SX * a.SY;
    const long lane_elem = (long)by * a.SX + (cvalid ? col0 + c : 0) + (long)g * zs;
    const ColPtr<S, OFF32> p(reinterpret_cast<S *>(a.data), 16 * zs, lane_elem);
#pragma unroll
    for (int q = 0; q < J; q++) {
        const int r = g + 16 * q;
        R v = (R)0;
        if (cvalid && (FULL || r < n)) v = (R)*p.at(q);
        tile[lrow(r) * (CPB + 1) + c] = v;
    }
}

template <typename R, int J, bool FULL, bool OFF32, int CPB, typename LRow, typename S = R>
__device__ __forceinline__ void ztile_store(const R *tile, const ZArgs &a, int n, LRow lrow, int bx, int by)
{
    const int tid = threadIdx.x;
    const int c = tid % CPB, g = tid / CPB;
    if (a.tiled) {
        if (a.tile_w == 0 || a.tile_w == CPB) {
            S *__restrict__ dst = reinterpret_cast<S *>(a.data) + ((long)bx * a.SY + by) * (a.tile_rows ? a.tile_rows : 16 * J) * CPB;
#pragma unroll
            for (int q = 0; q < J; q++) dst[q * 16 * CPB + tid] = (S)tile[lrow(g + 16 * q) * (CPB + 1) + c];
        } else {
            const int W = a.tile_w, sub = W / CPB;
            S *__restrict__ dst = reinterpret_cast<S *>(a.data) + ((long)(bx / sub) * a.SY + by) * (a.tile_rows ? a.tile_rows : 16 * J) * W +
                                  (bx % sub) * CPB + g * W + c;
#pragma unroll
            for (int q = 0; q < J; q++) dst[(long)q * 16 * W] = (S)tile[lrow(g + 16 * q) * (CPB + 1) + c];
        }
        return;
    }
    const int col0 = bx * CPB;
    const bool cvalid = col0 + c < a.SX;
    const long zs = (long)a.SX * a.SY;
    const long lane_elem = (long)by * a.SX + (cvalid ? col0 + c : 0) + (long)g * zs;
    const ColPtr<S, OFF32> p(reinterpret_cast<S *>(a.data), 16 * zs, lane_elem);
#pragma unroll
    for (int q = 0; q < J; q++) {
        const int r = g + 16 * q;
        if (cvalid && (FULL || r < n)) *p.at(q) = (S)tile[lrow(r) * (CPB + 1) + c];
    }
}

// LDS rows of the two layouts: cyclic (lane t owns rows t + 16 j) keeps tile order;
// blocked (lane t owns rows tJ .. tJ+J-1) gives each lane block an odd row stride LJ.
template <int J>
constexpr int blocked_lj() { return (J % 2 == 1) ? J : J + 1; }

// columns per block of the PCR z solve: 128-byte tile rows
template <typename R>
constexpr int zsolve_cpb() { return (int)(128 / sizeof(R)); }

template <typename R, int J, int CPB>
constexpr size_t ztile_bytes_cyclic() { return (size_t)16 * J * (CPB + 1) * sizeof(R); }
template <typename R, int J, int CPB>
constexpr size_t ztile_bytes_blocked() { return (size_t)16 * blocked_lj<J>() * (CPB + 1) * sizeof(R); }

// FULL: nz == 16 * J (no padding rows, no per-row bounds checks).  S: element type in
// memory (FP32 data solved in FP64 when R = double, S = float).  FIX: solve and write back
// only the near-singular columns (|eig+2| < tau, not pinned) that the FP32 solve left
// untouched; the tile's other columns are stored back unchanged.
template <typename R, int J, bool FULL, bool OFF32, int CPB = 16, typename S = R, bool FIX = false>
__global__ void __launch_bounds__(16 * CPB) k_zsolve(ZArgs a0)
{
    constexpr bool ALL = (J <= 3);
    constexpr int NEX = ALL ? J : 3;
    // FP64 columns beyond 32 rows per lane (nz > 512): the Thomas factors are kept in this
    // lane's own tile slots, free once its rows are in registers, instead of J more register
    // pairs.  The launch requires every column to be written back (tau = 0), since the slots
    // no longer hold the input afterwards.
    constexpr bool CP_LDS = sizeof(R) * J > 256;
    ZArgs a = a0;
    int bx = blockIdx.x, by = blockIdx.y;
    if (CPB * sizeof(S) < 128 && !a0.items) {
        // rows shorter than a 128-byte line (8 FP64 columns): the two blocks sharing a line
        // run on one XCD, so the line is fetched into one L2
        const int lin = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
        bx = lin % gridDim.x;
        by = lin / gridDim.x;
    }
    if (a0.items) {
        const ZItem it = a0.items[blockIdx.x];
        const ZSolve sv = a0.solves[it.solve];
        a.data = reinterpret_cast<S *>(a0.data) + sv.offset;
        a.SX = sv.SX;
        a.SY = sv.SY;
        a.nz = sv.nz;
        a.zlo_dirichlet = sv.zlo_dirichlet;
        a.zhi_dirichlet = sv.zhi_dirichlet;
        a.Cx = sv.Cx;
        a.Cy = sv.Cy;
        a.lamx = sv.lamx;
        a.lamy = sv.lamy;
        bx = it.bx;
        by = it.by;
    }
    const int tid = threadIdx.x;
    const int t = tid & 15;
    const int c = tid >> 4;
    int sx = bx * CPB + c;
    const int sy = by;
    const bool valid = sx < a.SX;
    if (!valid) sx = a.SX - 1;
    const int n = FULL ? 16 * J : a.nz;
    R *__restrict__ data = reinterpret_cast<R *>(a.data);
    const long zs = (long)a.SX * a.SY;
    const long col = (long)sy * a.SX + sx;

    // eigenvalue of this (kx, ky) column, src/pencilDcmp.cpp:4393-4417, without contraction
    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    bool pin = fabs(eig + 2.) < 1.e-10;
    if (pin && a.zero_mode == 1 && (a.zlo_dirichlet || a.zhi_dirichlet)) pin = false;
    const double b0d = pin ? 1.0 : (a.zlo_dirichlet ? eig - 1. : eig + 1.);
    const double c0d = pin ? 0.0 : 1.0;
    const double bnd = a.zhi_dirichlet ? eig - 1. : eig + 1.;

    auto row_coef = [&](int i) -> Row3<R> {
        if (i >= n) return {(R)0, (R)1, (R)0};
        if (i == 0) return {(R)0, (R)b0d, (R)c0d};
        if (i == n - 1) return {(R)1, (R)bnd, (R)0};
        return {(R)1, (R)eig, (R)1};
    };

    extern __shared__ __align__(16) unsigned char zsm[];
    R *tile = reinterpret_cast<R *>(zsm);
    auto lrow = [](int r) { return r; };
    ztile_load<R, J, FULL, OFF32, CPB, decltype(lrow), S>(tile, a, n, lrow, bx, by);
    __syncthreads();
    R d[J];
#pragma unroll
    for (int j = 0; j < J; j++) d[j] = tile[(t + 16 * j) * (CPB + 1) + c];
    Row3<R> ex[NEX];
    if constexpr (ALL) {
#pragma unroll
        for (int j = 0; j < J; j++) ex[j] = row_coef(t + 16 * j);
    } else {
        ex[0] = row_coef(t);
        ex[1] = row_coef(t + 16 * (J - 2));
        ex[2] = row_coef(t + 16 * (J - 1));
    }
    R ta = (R)1, tb = (R)eig;

    pcr_step<R, J, ALL, 1>(d, ex, ta, tb, t);
    pcr_step<R, J, ALL, 2>(d, ex, ta, tb, t);
    pcr_step<R, J, ALL, 4>(d, ex, ta, tb, t);
    pcr_step<R, J, ALL, 8>(d, ex, ta, tb, t);

    // Thomas on this lane's decoupled system (rows t, t+16, ...)
    auto coef = [&](int j) -> Row3<R> {
        if constexpr (ALL) return ex[j];
        else {
            if (j == 0) return ex[0];
            if (j == J - 2) return ex[1];
            if (j == J - 1) return ex[2];
            return Row3<R>{ta, tb, ta};
        }
    };
    R cp[CP_LDS ? 1 : J];
    R *const cps = tile + t * (CPB + 1) + c;  // this lane's slot of row j: cps[j * 16 * (CPB + 1)]
    R cprev;
    {
        Row3<R> r0 = coef(0);
        const R inv = rcp_nr(r0.b);
        cprev = r0.c * inv;
        if constexpr (CP_LDS) cps[0] = cprev;
        else cp[0] = cprev;
        d[0] = d[0] * inv;
    }
#pragma unroll
    for (int j = 1; j < J; j++) {
        Row3<R> r = coef(j);
        const R inv = rcp_nr(r.b - r.a * cprev);
        cprev = r.c * inv;
        if constexpr (CP_LDS) cps[j * 16 * (CPB + 1)] = cprev;
        else cp[j] = cprev;
        d[j] = (d[j] - r.a * d[j - 1]) * inv;
        // factors to LDS: the row update is pinned here (otherwise it sinks to the backward
        // sweep and every pivot stays live until then)
        if constexpr (CP_LDS) asm volatile("" : "+v"(d[j]));
    }
#pragma unroll
    for (int j = J - 2; j >= 0; j--) {
        R cj;
        if constexpr (CP_LDS) cj = cps[j * 16 * (CPB + 1)];
        else cj = cp[j];
        d[j] = d[j] - cj * d[j + 1];
        // factors from LDS: loaded a few rows ahead at most, not all J beside the J rows
        if constexpr (CP_LDS) {
            if (j % 4 == 0) asm volatile("" : "+v"(d[j])::"memory");
        }
    }

    const bool near = !pin && fabs(eig + 2.) < a.tau;
    const bool skip = FIX ? !near : near;
    const bool zero = pin && a.pinned_zero;
    if (!skip) {
#pragma unroll
        for (int j = 0; j < J; j++) tile[(t + 16 * j) * (CPB + 1) + c] = zero ? (R)0 : d[j];
    }
    __syncthreads();
    ztile_store<R, J, FULL, OFF32, CPB, decltype(lrow), S>(tile, a, n, lrow, bx, by);
    (void)valid;
    (void)data;
}

// ---------------------------------------------------------------------------------------
// Partitioned Thomas with cross-lane scans (second z solver).
//
// Lane t of a column owns the contiguous rows tJ .. tJ+J-1.  LU without pivoting:
//   m_i = a_i / u_{i-1},  u_i = b_i - m_i c_{i-1},  y_i = d_i - m_i y_{i-1},
//   x_i = (y_i - c_i x_{i+1}) / u_i.
// The pivot entering a lane, u_{tJ-1}, follows from u_1 by the Moebius map
// u -> b - 1/u, i.e. the vector (p, q) ~ u = p/q multiplied by G^(tJ-2), G = [[b,-1],[1,0]],
// computed by binary powering with exact power-of-two rescaling.  The forward and backward
// sweeps run locally per lane from zero boundary values; their affine dependence on the
// neighbouring lane's value is stitched by Kogge-Stone scans over the 16 lanes (DPP), then
// one fix-up pass per direction.
// ---------------------------------------------------------------------------------------
struct Mat2 {
    double a, b, c, d;
};

__device__ __forceinline__ double pow2_scale(double x, int e) { return __builtin_amdgcn_ldexp(x, e); }

__device__ __forceinline__ Mat2 mat_sq_norm(Mat2 m)
{
    Mat2 r = {m.a * m.a + m.b * m.c, m.a * m.b + m.b * m.d, m.c * m.a + m.d * m.c, m.c * m.b + m.d * m.d};
    const double mx = fmax(fmax(fabs(r.a), fabs(r.b)), fmax(fabs(r.c), fabs(r.d)));
    const int e = -__builtin_amdgcn_frexp_exp(mx);
    return {pow2_scale(r.a, e), pow2_scale(r.b, e), pow2_scale(r.c, e), pow2_scale(r.d, e)};
}

constexpr int bits_for(int v)
{
    int b = 0;
    while ((1 << b) <= v) b++;
    return b;
}

template <int S, typename R>
__device__ __forceinline__ R from_lower(R x) { return rot_from_lower<S>(x); }
template <int S, typename R>
__device__ __forceinline__ R from_upper(R x) { return rot_from_upper<S>(x); }

template <typename R, int J, bool FULL, bool OFF32, int CPB = 16>
__global__ void __launch_bounds__(16 * CPB) k_zsolve_scan(ZArgs a)
{
    const int tid = threadIdx.x;
    const int t = tid & 15;
    const int cidx = tid >> 4;
    int sx = blockIdx.x * CPB + cidx;
    const int sy = blockIdx.y;
    const bool valid = sx < a.SX;
    if (!valid) sx = a.SX - 1;
    const int n = FULL ? 16 * J : a.nz;
    R *__restrict__ data = reinterpret_cast<R *>(a.data);
    const long zs = (long)a.SX * a.SY;
    const long col = (long)sy * a.SX + sx;

    const double eig = __dadd_rn(__dadd_rn(-2.0, __dmul_rn(a.lamx[sx], a.Cx)), __dmul_rn(a.lamy[sy], a.Cy));
    bool pin = fabs(eig + 2.) < 1.e-10;
    if (pin && a.zero_mode == 1 && (a.zlo_dirichlet || a.zhi_dirichlet)) pin = false;
    const double b0 = pin ? 1.0 : (a.zlo_dirichlet ? eig - 1. : eig + 1.);
    const double c0 = pin ? 0.0 : 1.0;
    const double bn = a.zhi_dirichlet ? eig - 1. : eig + 1.;

    const int i0 = t * J;  // first row of this lane
    extern __shared__ __align__(16) unsigned char zsm[];
    R *tile = reinterpret_cast<R *>(zsm);
    constexpr int LJ = blocked_lj<J>();
    auto lrow = [](int r) { return (r / J) * LJ + (r % J); };
    ztile_load<R, J, FULL, OFF32, CPB>(tile, a, n, lrow, blockIdx.x, blockIdx.y);
    __syncthreads();
    R d[J];
#pragma unroll
    for (int j = 0; j < J; j++) d[j] = tile[(t * LJ + j) * (CPB + 1) + cidx];

    // ---- pivot entering the lane: 1/u_{i0-1} ----
    double iu_in = 0.0;
    if (i0 >= 1 && i0 <= n - 1) {
        if (i0 == 1) {
            iu_in = 1.0 / b0;
        } else {
            // (p, q) for u_1 = b - c0/b0, then G^(i0-2)
            double p = eig * b0 - c0, q = b0;
            Mat2 Q = {eig, -1.0, 1.0, 0.0};
            const int e = i0 - 2;
            constexpr int KB = bits_for(16 * J);
#pragma unroll
            for (int k = 0; k < KB; k++) {
                if ((e >> k) & 1) {
                    const double np = Q.a * p + Q.b * q, nq = Q.c * p + Q.d * q;
                    const int sc = -__builtin_amdgcn_frexp_exp(fmax(fabs(np), fabs(nq)));
                    p = pow2_scale(np, sc);
                    q = pow2_scale(nq, sc);
                }
                if (k + 1 < KB) Q = mat_sq_norm(Q);
            }
            iu_in = q / p;
        }
    }

    // row coefficients a_i (sub), b_i (diag), c_i (super) of global row i = i0 + j.
    // FULL: row 0 is (lane 0, j = 0) and row n-1 is (lane 15, j = J-1).
    const bool first = (t == 0), lastl = (t == 15);
    auto is_row0 = [&](int j) -> bool { return FULL ? (first && j == 0) : (i0 + j == 0); };
    auto is_rowN = [&](int j) -> bool { return FULL ? (lastl && j == J - 1) : (i0 + j == n - 1); };
    auto is_pad = [&](int j) -> bool { return FULL ? false : (i0 + j >= n); };
    auto ai = [&](int j) -> R { return (is_row0(j) || is_pad(j)) ? (R)0 : (R)1; };
    auto bi = [&](int j) -> R {
        if (is_row0(j)) return (R)b0;
        if (is_rowN(j)) return (R)bn;
        if (is_pad(j)) return (R)1;
        return (R)eig;
    };
    auto ci = [&](int j) -> R {
        if (is_row0(j)) return (R)c0;
        if (is_rowN(j) || is_pad(j)) return (R)0;
        return (R)1;
    };
    // c_{i-1} for the row above global row i0 + j (j = 0: the previous lane's last row)
    auto cprev = [&](int j) -> R {
        if (j > 0) return ci(j - 1);
        if (i0 == 0) return (R)0;
        if (i0 == 1) return (R)c0;
        return (FULL || i0 - 1 < n - 1) ? (R)1 : (R)0;
    };

    // Every pass after a scan is a true recurrence on the scanned boundary value, so no
    // scan-independent per-row product is left for the compiler to hoist: only d[] and
    // iu[] (1/u) stay live across passes.

    // ---- pass A: pivots 1/u_i, and the lane's forward map y_end = yh + piA * y_{i0-1} ----
    R iu[J];
    R yh = (R)0, piA = (R)1;
    {
        R iup = (R)iu_in;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const R m = ai(j) * iup;
            const R u = bi(j) - m * cprev(j);
            const R iuj = rcp_nr(u);
            yh = d[j] - m * yh;
            piA = piA * (-m);
            iu[j] = iuj;
            iup = iuj;
        }
    }

    // ---- forward scan over lanes: y_end(t) = B + A * y_end(t-1) ----
    R A = piA, B = yh;
    {
#define PP_FWD_STEP(S)                          \
    {                                           \
        const R Al = from_lower<S>(A);          \
        const R Bl = from_lower<S>(B);          \
        if (t >= S) {                           \
            B = B + A * Bl;                     \
            A = A * Al;                         \
        }                                       \
    }
        PP_FWD_STEP(1) PP_FWD_STEP(2) PP_FWD_STEP(4) PP_FWD_STEP(8)
#undef PP_FWD_STEP
    }
    const R yin_r = from_lower<1>(B);
    const R y_in = (t >= 1) ? yin_r : (R)0;

    // ---- pass B: forward sweep y_i = d_i - m_i y_{i-1} from the exact y_{i0-1} ----
    {
        R yprev = y_in, iup = (R)iu_in;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const R m = ai(j) * iup;
            const R y = d[j] - m * yprev;
            d[j] = y;
            yprev = y;
            iup = iu[j];
        }
    }

    // ---- pass B2: the lane's backward map x_start = xh + sig * x_{i0+J} ----
    R xh = (R)0, sig = (R)1;
#pragma unroll
    for (int j = J - 1; j >= 0; j--) {
        const R cc = ci(j);
        xh = (d[j] - cc * xh) * iu[j];
        sig = sig * (-cc * iu[j]);
    }

    // ---- backward scan over lanes: x_start(t) = B + A * x_start(t+1) ----
    A = sig;
    B = xh;
    {
#define PP_BWD_STEP(S)                          \
    {                                           \
        const R Au = from_upper<S>(A);          \
        const R Bu = from_upper<S>(B);          \
        if (t + S < 16) {                       \
            B = B + A * Bu;                     \
            A = A * Au;                         \
        }                                       \
    }
        PP_BWD_STEP(1) PP_BWD_STEP(2) PP_BWD_STEP(4) PP_BWD_STEP(8)
#undef PP_BWD_STEP
    }
    const R xout_r = from_upper<1>(B);
    const R x_out = (t < 15) ? xout_r : (R)0;

    // ---- pass C: backward sweep x_i = (y_i - c_i x_{i+1}) / u_i from the exact x_{i0+J} ----
    {
        R xn = x_out;
#pragma unroll
        for (int j = J - 1; j >= 0; j--) {
            const R x = (d[j] - ci(j) * xn) * iu[j];
            d[j] = x;
            xn = x;
        }
    }

    const bool skip = !pin && fabs(eig + 2.) < a.tau;
    const bool zero = pin && a.pinned_zero;
    if (!skip) {
#pragma unroll
        for (int j = 0; j < J; j++) tile[(t * LJ + j) * (CPB + 1) + cidx] = zero ? (R)0 : d[j];
    }
    __syncthreads();
    ztile_store<R, J, FULL, OFF32, CPB>(tile, a, n, lrow, blockIdx.x, blockIdx.y);
    (void)valid;
    (void)data;
    (void)col;
}

}  // namespace pp
