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


namespace pp {

// ---------------------------------------------------------------------------------------
// z solve
// ---------------------------------------------------------------------------------------
struct ZArgs {
    void *data;  // [nz][SY][SX], in place
    int SX, SY, nz;
    const double *lamx;  // -2 + 2 cos(...) per x mode (SX)
    const double *lamy;  // per y mode (SY)
    double Cx, Cy;
    int zlo_dirichlet, zhi_dirichlet;  // 1 = D, 0 = N
    int zero_mode;                      // 0: reference row-0 rule; 1: only when singular (z NN)
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

// FULL: nz == 16 * J (no padding rows, no per-row bounds checks)
template <typename R, int J, bool FULL>
__global__ void __launch_bounds__(256) k_zsolve(ZArgs a)
{
    constexpr bool ALL = (J <= 3);
    constexpr int NEX = ALL ? J : 3;
    const int tid = threadIdx.x;
    const int t = tid & 15;
    const int c = tid >> 4;
    int sx = blockIdx.x * 16 + c;
    const int sy = blockIdx.y;
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

    R d[J];
    const R *src = data + col + (long)t * zs;
    const long zs16 = 16 * zs;
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int i = t + 16 * j;
        d[j] = (FULL || i < n) ? src[j * zs16] : (R)0;
    }
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
    R cp[J];
    {
        Row3<R> r0 = coef(0);
        const R inv = rcp_nr(r0.b);
        cp[0] = r0.c * inv;
        d[0] = d[0] * inv;
    }
#pragma unroll
    for (int j = 1; j < J; j++) {
        Row3<R> r = coef(j);
        const R inv = rcp_nr(r.b - r.a * cp[j - 1]);
        cp[j] = r.c * inv;
        d[j] = (d[j] - r.a * d[j - 1]) * inv;
    }
#pragma unroll
    for (int j = J - 2; j >= 0; j--) d[j] = d[j] - cp[j] * d[j + 1];

    if (valid) {
        R *dst = data + col + (long)t * zs;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int i = t + 16 * j;
            if (FULL || i < n) dst[j * zs16] = d[j];
        }
    }
}

// lane-rotation self test: out[lane] = value rotated from lane (t - S) for S = 1,2,4,8
__global__ void k_selftest_rot(int *out)
{
    const int lane = threadIdx.x;
    const float x = (float)lane;
    out[lane + 0 * 64] = (int)rot_from_lower<1>(x);
    out[lane + 1 * 64] = (int)rot_from_lower<2>(x);
    out[lane + 2 * 64] = (int)rot_from_lower<4>(x);
    out[lane + 3 * 64] = (int)rot_from_lower<8>(x);
    out[lane + 4 * 64] = (int)rot_from_upper<1>((double)lane);
    out[lane + 5 * 64] = (int)rot_from_upper<8>((double)lane);
}

}  // namespace pp
