// pp_kernels.h -- HIP kernels of the solve path for CDNA4 (gfx950).
//
//  * k_line   : batched real-to-real line transform along x (contiguous lines) or y
//               (strided columns), with the DCT/DST/R2HC pre- and post-processing, the
//               RHS scaling, the Dirichlet boundary term and the final 1/(Nx*Ny) fused
//               into the load and store phases.  Replaces SignalProc's pre/post steps
//               (src/signalProc.cpp:32-265), the cuFFT calls (src/poissonGPU.cpp:5-98),
//               assignRhs's dz^2 (src/pencilDcmp.cpp:5591-5647), modifyRhsDirichlet
//               (4623-4835), the x<->y plane transposes (2279-2515) and rescale (3503-3531).
//  * k_zsolve : batched tridiagonal z solve per (kx, ky) column, PCR across 16 lanes (DPP
//               row rotations) then Thomas per lane.  Replaces solveThmBatch +
//               fillInArrayContig/Back + thomasLowMem (src/pencilDcmp.cpp:5062-5234,
//               src/triDiag.cpp:1629-1822).
//
// All transforms keep a real line of length N real: the periodic axis uses the
// half-complex R2HC layout, so the spectral field has the size of the physical field and
// the z solve of column (sx, sy) uses the reference eigenvalue with j = sx, i = sy.
#pragma once
#include <hip/hip_runtime.h>

#include "pp_fft.h"

namespace pp {

enum AxisKind : int { KIND_DD = 0, KIND_NN = 1, KIND_PP = 2, KIND_ND = 3, KIND_DN = 4 };

struct LineArgs {
    const void *in;
    void *out;
    long nlines;  // x: number of lines (ny*nz); y: columns per plane (nx)
    int ncols;    // y: element stride (= nx); x: unused
    int nplanes;  // y: nz
    int kind;     // AxisKind of this axis
    const void *tw, *A, *B, *C, *D;  // twiddle tables, compute precision, complex
    double in_scale;
    double out_scale;
    // Dirichlet boundary term (forward x only): dtab = [ex(nx) ey(ny) ez(nz) exf0 exf1 eyf0
    // eyf1 ezf0 ezf1], dmask bit f set when face f (xmin,xmax,ymin,ymax,zmin,zmax) is D.
    const double *dtab;
    int dmask;
    int nx, ny, nz;
};

template <typename T>
__device__ __forceinline__ double to_d(T v) { return (double)v; }

// position in the line's real LDS view where input element n is placed (forward DCT/DST:
// Makhoul even/odd reorder; DCT-IV/DST-IV: interleave x_{2n} / x_{N-1-2n})
template <int N>
__device__ __forceinline__ int load_pos(int kind, bool fwd, int n, bool &neg)
{
    neg = false;
    if (kind == KIND_ND || kind == KIND_DN) {
        int a = (kind == KIND_DN) ? (N - 1 - n) : n;
        return (a & 1) ? (N - a) : a;
    }
    if (fwd && (kind == KIND_NN || kind == KIND_DD)) {
        if (kind == KIND_DD) neg = (n & 1);
        return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
    }
    return n;
}

template <int N>
__device__ __forceinline__ int store_pos(int kind, bool fwd, int n, bool &neg)
{
    neg = false;
    if (!fwd && (kind == KIND_NN || kind == KIND_DD)) {
        if (kind == KIND_DD) neg = (n & 1);
        return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
    }
    return n;
}

template <typename TIn, typename TOut, typename R, int M, int T, int LINES, int AX, int FWD>
__global__ void __launch_bounds__(T *LINES) k_line(LineArgs a)
{
    constexpr int N = 2 * M;
    constexpr int E = M / T;
    constexpr int LS = M + 1;  // LDS line stride (complex elements), padded
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_raw);

    const int tid = threadIdx.x;
    int l, t;
    if constexpr (AX == 0) {
        t = tid % T;
        l = tid / T;
    } else {
        l = tid % LINES;
        t = tid / LINES;
    }
    cpx<R> *L = lds + l * LS;
    R *Lr = reinterpret_cast<R *>(L);

    const TIn *__restrict__ in = reinterpret_cast<const TIn *>(a.in);
    TOut *__restrict__ out = reinterpret_cast<TOut *>(a.out);
    const cpx<R> *__restrict__ tw = reinterpret_cast<const cpx<R> *>(a.tw);
    const cpx<R> *__restrict__ TA = reinterpret_cast<const cpx<R> *>(a.A);
    const cpx<R> *__restrict__ TB = reinterpret_cast<const cpx<R> *>(a.B);
    const cpx<R> *__restrict__ TC = reinterpret_cast<const cpx<R> *>(a.C);
    const cpx<R> *__restrict__ TD = reinterpret_cast<const cpx<R> *>(a.D);
    const int kind = a.kind;
    constexpr bool fwd = FWD != 0;

    long base, stride;
    bool valid;
    int ly = 0, lz = 0;
    if constexpr (AX == 0) {
        long line = (long)blockIdx.x * LINES + l;
        valid = line < a.nlines;
        if (!valid) line = 0;
        base = line * N;
        stride = 1;
        ly = (int)(line % a.ny);
        lz = (int)(line / a.ny);
    } else {
        long col = (long)blockIdx.x * LINES + l;
        valid = col < a.ncols;
        if (!valid) col = 0;
        base = (long)blockIdx.y * N * a.ncols + col;
        stride = a.ncols;
    }

    // ---- load (coalesced) + fused input scaling / Dirichlet term, scatter into LDS ----
#pragma unroll
    for (int c = 0; c < 2 * E; c++) {
        const int n = t + T * c;
        double v = 0.0;
        if (valid) {
            v = to_d(in[base + (long)n * stride]);
            if constexpr (AX == 0 && FWD) {
                v = v * a.in_scale;  // a.in_scale = c3 (applied twice: rhs*c3*c3)
                v = v * a.in_scale;
                if (a.dmask) {
                    const double *ex = a.dtab, *ey = a.dtab + a.nx, *ez = a.dtab + a.nx + a.ny;
                    const double *ef = a.dtab + a.nx + a.ny + a.nz;
                    if ((a.dmask & 16) && lz == 0) v = v - 2. * ex[n] * ey[ly] * ef[4];
                    if ((a.dmask & 32) && lz == a.nz - 1) v = v - 2 * ex[n] * ey[ly] * ef[5];
                    if ((a.dmask & 1) && n == 0) v = v - 2. * ef[0] * ey[ly] * ez[lz];
                    if ((a.dmask & 2) && n == N - 1) v = v - 2. * ef[1] * ey[ly] * ez[lz];
                    if ((a.dmask & 4) && ly == 0) v = v - 2. * ex[n] * ef[2] * ez[lz];
                    if ((a.dmask & 8) && ly == a.ny - 1) v = v - 2. * ex[n] * ef[3] * ez[lz];
                }
            } else {
                v = v * a.in_scale;
            }
        }
        bool neg;
        const int pos = load_pos<N>(kind, fwd, n, neg);
        Lr[pos] = neg ? (R)(-v) : (R)v;
    }
    __syncthreads();

    // ---- pre-processing: FFT input in registers, layout k = t + T*m ----
    cpx<R> v[E];
    if (kind == KIND_ND || kind == KIND_DN) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            v[m] = c_mul(L[k], TC[k]);
        }
    } else if (fwd) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = L[t + T * m];
    } else if (kind == KIND_PP) {
        // half-complex -> packed half-length spectrum: Z_k = (X_k + conj X_{M-k})
        //                                              + i e^{+2 pi i k/N} (X_k - conj X_{M-k})
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            const int k2 = M - k;
            cpx<R> X1 = {Lr[k], (k == 0) ? (R)0 : Lr[N - k]};
            cpx<R> X2 = {Lr[k2], (k2 == M) ? (R)0 : Lr[N - k2]};
            cpx<R> s = {X1.x + X2.x, X1.y - X2.y};
            cpx<R> d = {X1.x - X2.x, X1.y + X2.y};
            cpx<R> w = c_conj(TA[k]);
            cpx<R> q = c_mul(w, d);
            v[m] = {s.x - q.y, s.y + q.x};
        }
    } else {
        // DCT-III/2 (NN) or DST-III/2 (DD, input reversed):
        // V_k = 1/2 e^{i pi k/(2N)} (X_k - i X_{N-k}),  X_N = 0
        // Z_k = (V_k + conj V_{M-k}) + i e^{2 pi i k/N} (V_k - conj V_{M-k})
        const bool rev = (kind == KIND_DD);
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            const int k2 = M - k;
            auto X = [&](int q) -> R {
                if (q >= N) return (R)0;
                return rev ? Lr[N - 1 - q] : Lr[q];
            };
            cpx<R> V1, V2;
            {
                cpx<R> u = {X(k), -X(N - k)};
                V1 = c_scale(c_mulc(u, TB[k]), (R)0.5);
            }
            {
                cpx<R> u = {X(k2), -X(N - k2)};
                V2 = c_scale(c_mulc(u, TB[k2]), (R)0.5);
            }
            cpx<R> s = {V1.x + V2.x, V1.y - V2.y};
            cpx<R> d = {V1.x - V2.x, V1.y + V2.y};
            cpx<R> q = c_mul(c_conj(TA[k]), d);
            v[m] = {s.x - q.y, s.y + q.x};
        }
    }
    __syncthreads();

    // ---- complex FFT of length M ----
    const bool inv_fft = (!fwd) && (kind == KIND_PP || kind == KIND_NN || kind == KIND_DD);
    if (inv_fft) line_fft<R, M, T, +1>(v, L, t, tw);
    else line_fft<R, M, T, -1>(v, L, t, tw);

    // ---- post-processing into the line's real LDS view ----
    if (fwd && (kind == KIND_PP || kind == KIND_NN || kind == KIND_DD)) {
#pragma unroll
        for (int m = 0; m < E; m++) L[t + T * m] = v[m];
        __syncthreads();
        R o1[E], o2[E];
        int p1[E], p2[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            cpx<R> Zk = L[k];
            cpx<R> Zc = c_conj(L[k == 0 ? 0 : M - k]);
            // X_k = 1/2 (Zk + Zc) - 1/2 i e^{-2 pi i k/N} (Zk - Zc)
            cpx<R> s = c_scale(c_add(Zk, Zc), (R)0.5);
            cpx<R> q = c_mul(TA[k], c_scale(c_sub(Zk, Zc), (R)0.5));
            cpx<R> X = {s.x + q.y, s.y - q.x};
            if (kind == KIND_PP) {
                if (k == 0) {
                    o1[m] = Zk.x + Zk.y;
                    p1[m] = 0;
                    o2[m] = Zk.x - Zk.y;
                    p2[m] = M;
                } else {
                    o1[m] = X.x;
                    p1[m] = k;
                    o2[m] = X.y;
                    p2[m] = N - k;
                }
            } else {
                // DCT-II: X_k = 2 Re(e^{-i pi k/(2N)} V_k), X_{N-k} = -2 Im(...)
                if (k == 0) {
                    o1[m] = (R)2 * X.x;  // V_0 real
                    p1[m] = 0;
                    R VM = Zk.x - Zk.y;  // V_M (real)
                    o2[m] = (R)2 * TB[M].x * VM;
                    p2[m] = M;
                } else {
                    cpx<R> w = c_mul(TB[k], X);
                    o1[m] = (R)2 * w.x;
                    p1[m] = k;
                    o2[m] = (R)-2 * w.y;
                    p2[m] = N - k;
                }
                if (kind == KIND_DD) {
                    p1[m] = N - 1 - p1[m];
                    p2[m] = N - 1 - p2[m];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < E; m++) {
            Lr[p1[m]] = o1[m];
            Lr[p2[m]] = o2[m];
        }
    } else if (kind == KIND_ND || kind == KIND_DN) {
        const R f = fwd ? (R)2 : (R)1;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            cpx<R> y = c_mul(v[m], TD[k]);
            Lr[2 * k] = f * y.x;
            Lr[N - 1 - 2 * k] = (kind == KIND_DN) ? f * y.y : -f * y.y;
        }
    } else {
        // inverse PP / NN / DD: z_m holds (x_{2m}, x_{2m+1}) or Makhoul-ordered v
#pragma unroll
        for (int m = 0; m < E; m++) L[t + T * m] = v[m];
    }
    __syncthreads();

    // ---- store (coalesced) with fused output scaling ----
    if (valid) {
#pragma unroll
        for (int c = 0; c < 2 * E; c++) {
            const int n = t + T * c;
            bool neg;
            const int pos = store_pos<N>(kind, fwd, n, neg);
            double val = (double)Lr[pos];
            if (neg) val = -val;
            val = a.out_scale * val;
            out[base + (long)n * stride] = (TOut)val;
        }
    }
}

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

// value of x from lane (t - S) mod 16 of this lane's 16-lane row
template <int S>
__device__ __forceinline__ float rot_from_lower(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x120 + S, 0xF, 0xF, false));
}
template <int S>
__device__ __forceinline__ double rot_from_lower(double x)
{
    int lo = __double2loint(x), hi = __double2hiint(x);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x120 + S, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x120 + S, 0xF, 0xF, false);
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
    const R k1 = me.a / L.b;
    const R k2 = me.c / Rr.b;
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
        const R kt = ta / tb;
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

template <typename R, int J>
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
    const int n = a.nz;
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
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int i = t + 16 * j;
        d[j] = (i < n) ? data[(long)i * zs + col] : (R)0;
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
        const R inv = (R)1 / r0.b;
        cp[0] = r0.c * inv;
        d[0] = d[0] * inv;
    }
#pragma unroll
    for (int j = 1; j < J; j++) {
        Row3<R> r = coef(j);
        const R inv = (R)1 / (r.b - r.a * cp[j - 1]);
        cp[j] = r.c * inv;
        d[j] = (d[j] - r.a * d[j - 1]) * inv;
    }
#pragma unroll
    for (int j = J - 2; j >= 0; j--) d[j] = d[j] - cp[j] * d[j + 1];

    if (valid) {
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int i = t + 16 * j;
            if (i < n) data[(long)i * zs + col] = d[j];
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
