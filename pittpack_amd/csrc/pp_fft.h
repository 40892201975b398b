// pp_fft.h -- line FFT engine for CDNA4 (gfx950): Stockham autosort, LDS exchanges.
//
// A complex line of length M is processed by T threads; thread t holds the E = M/T
// elements with index t + T*m (m = 0..E-1) in registers.  That layout is both the input
// layout of the first pass (so the first pass needs no exchange) and the output layout of
// the last pass (so results stay in registers).  Passes in between exchange through the
// line's LDS slice.  Radices are picked at compile time from {16, 8, 4, 2, 3, 5, 7} so that
// each divides E; every butterfly is unrolled in registers.
//
// Replaces the cuFFT/FFTW Z2Z calls of the reference (src/poissonGPU.cpp:5-98,
// src/poissonCPU.cpp:40-234): unnormalised, FORWARD = exp(-2 pi i jk/M), inverse = exp(+...).
#pragma once
#include <hip/hip_runtime.h>

namespace pp {

template <typename R>
struct cpx {
    R x, y;
};

template <typename R>
__device__ __forceinline__ cpx<R> c_add(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> c_sub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> c_mul(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename R>
__device__ __forceinline__ cpx<R> c_mulc(cpx<R> a, cpx<R> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename R>
__device__ __forceinline__ cpx<R> c_conj(cpx<R> a) { return {a.x, -a.y}; }
template <typename R>
__device__ __forceinline__ cpx<R> c_scale(cpx<R> a, R s) { return {a.x * s, a.y * s}; }
// multiply by -i (DIR = -1, forward) or +i (DIR = +1, inverse)
template <int DIR, typename R>
__device__ __forceinline__ cpx<R> c_rot(cpx<R> a)
{
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// ---------------------------------------------------------------------------------------
// Compile-time radix plan
// ---------------------------------------------------------------------------------------
struct RadixPlan {
    int n;
    int r[24];
};

constexpr RadixPlan make_plan(int M, int E)
{
    RadixPlan p{};
    p.n = 0;
    int rem = M;
    const int cand[7] = {16, 8, 4, 2, 3, 5, 7};
    while (rem > 1 && p.n < 24) {
        int pick = 0;
        for (int c : cand) {
            if (rem % c == 0 && E % c == 0) {
                pick = c;
                break;
            }
        }
        if (pick == 0) {
            p.n = -1;  // no valid plan
            return p;
        }
        p.r[p.n++] = pick;
        rem /= pick;
    }
    return p;
}

constexpr bool plan_ok(int M, int E) { return M % E == 0 && make_plan(M, E).n >= 0; }

// ---------------------------------------------------------------------------------------
// Small DFTs (in place on u[0..R-1]); DIR = -1 forward, +1 inverse
// ---------------------------------------------------------------------------------------
template <int DIR, typename R>
__device__ __forceinline__ void dft2(cpx<R> *u)
{
    cpx<R> a = u[0], b = u[1];
    u[0] = c_add(a, b);
    u[1] = c_sub(a, b);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft4(cpx<R> *u)
{
    cpx<R> a0 = c_add(u[0], u[2]), a1 = c_sub(u[0], u[2]);
    cpx<R> a2 = c_add(u[1], u[3]), a3 = c_rot<DIR>(c_sub(u[1], u[3]));
    u[0] = c_add(a0, a2);
    u[2] = c_sub(a0, a2);
    u[1] = c_add(a1, a3);
    u[3] = c_sub(a1, a3);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft8(cpx<R> *u)
{
    const R h = (R)0.70710678118654752440084436210485;
    cpx<R> a[4], b[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        a[r] = c_add(u[r], u[r + 4]);
        b[r] = c_sub(u[r], u[r + 4]);
    }
    // b[r] *= W8^r  (W8 = exp(DIR * i pi/4))
    {
        cpx<R> t = b[1];
        // W8^1 = h*(1 + DIR*i)
        b[1] = DIR < 0 ? cpx<R>{h * (t.x + t.y), h * (t.y - t.x)} : cpx<R>{h * (t.x - t.y), h * (t.y + t.x)};
        b[2] = c_rot<DIR>(b[2]);
        t = b[3];
        // W8^3 = h*(-1 + DIR*i)
        b[3] = DIR < 0 ? cpx<R>{h * (t.y - t.x), -h * (t.x + t.y)} : cpx<R>{-h * (t.x + t.y), h * (t.x - t.y)};
    }
    dft4<DIR>(a);
    dft4<DIR>(b);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        u[2 * r] = a[r];
        u[2 * r + 1] = b[r];
    }
}

template <int DIR, typename R>
__device__ __forceinline__ void dft16(cpx<R> *u)
{
    // 4 x 4: n = n1 + 4 n2, k = k2 + 4 k1
    cpx<R> A[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) {
        cpx<R> v[4] = {u[n1], u[n1 + 4], u[n1 + 8], u[n1 + 12]};
        dft4<DIR>(v);
#pragma unroll
        for (int k2 = 0; k2 < 4; k2++) A[n1][k2] = v[k2];
    }
    const R c1 = (R)0.92387953251128675612818318939679;  // cos(pi/8)
    const R s1 = (R)0.38268343236508977172845998403040;  // sin(pi/8)
    const R h = (R)0.70710678118654752440084436210485;
    // twiddle A[n1][k2] *= W16^(n1*k2), W16 = exp(DIR * 2 pi i / 16)
    const R tc[10] = {(R)1, c1, h, s1, (R)0, -s1, -h, -c1, (R)-1, -c1};
    const R ts[10] = {(R)0, s1, h, c1, (R)1, c1, h, s1, (R)0, -s1};
#pragma unroll
    for (int n1 = 1; n1 < 4; n1++) {
#pragma unroll
        for (int k2 = 1; k2 < 4; k2++) {
            const int e = n1 * k2;  // 1..9
            cpx<R> w = {tc[e], (R)DIR * ts[e]};
            A[n1][k2] = c_mul(A[n1][k2], w);
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) {
        cpx<R> v[4] = {A[0][k2], A[1][k2], A[2][k2], A[3][k2]};
        dft4<DIR>(v);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) u[k2 + 4 * k1] = v[k1];
    }
}

template <int DIR, typename R>
__device__ __forceinline__ void dft3(cpx<R> *u)
{
    const R c = (R)-0.5;
    const R s = (R)DIR * (R)0.86602540378443864676372317075294;  // sin(2pi/3) * DIR
    cpx<R> a = u[0], b = u[1], d = u[2];
    cpx<R> t1 = c_add(b, d), t2 = c_sub(b, d);
    u[0] = c_add(a, t1);
    cpx<R> m = {a.x + c * t1.x, a.y + c * t1.y};
    cpx<R> q = {-s * t2.y, s * t2.x};  // s * i * t2
    u[1] = c_add(m, q);
    u[2] = c_sub(m, q);
}

template <int DIR, typename R>
__device__ __forceinline__ void dft5(cpx<R> *u)
{
    const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
    const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
    const R s1 = (R)DIR * (R)0.95105651629515357211643933337938;
    const R s2 = (R)DIR * (R)0.58778525229247312916870595463907;
    cpx<R> t1 = c_add(u[1], u[4]), t2 = c_add(u[2], u[3]);
    cpx<R> t3 = c_sub(u[1], u[4]), t4 = c_sub(u[2], u[3]);
    cpx<R> a = u[0];
    u[0] = c_add(a, c_add(t1, t2));
    cpx<R> m1 = {a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
    cpx<R> m2 = {a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
    // i*(s1 t3 + s2 t4),  i*(s2 t3 - s1 t4)
    cpx<R> n1 = {-(s1 * t3.y + s2 * t4.y), s1 * t3.x + s2 * t4.x};
    cpx<R> n2 = {-(s2 * t3.y - s1 * t4.y), s2 * t3.x - s1 * t4.x};
    u[1] = c_add(m1, n1);
    u[4] = c_sub(m1, n1);
    u[2] = c_add(m2, n2);
    u[3] = c_sub(m2, n2);
}

// 7-point DFT by symmetric pairs: with t_j = u_j + u_{7-j}, s_j = u_j - u_{7-j} (j = 1..3),
// X_0 = u_0 + t_1 + t_2 + t_3 and, for k = 1..3, a_k = u_0 + sum_j cos(2 pi jk/7) t_j,
// b_k = sum_j sin(2 pi jk/7) s_j, X_k = a_k + i DIR b_k, X_{7-k} = a_k - i DIR b_k
template <int DIR, typename R>
__device__ __forceinline__ void dft7(cpx<R> *u)
{
    const R c1 = (R)0.62348980185873353052500488400424;   // cos(2pi/7)
    const R c2 = (R)-0.22252093395631440428890256449679;  // cos(4pi/7)
    const R c3 = (R)-0.90096886790241912623610231950745;  // cos(6pi/7)
    const R s1 = (R)DIR * (R)0.78183148246802980870844452667406;  // sin(2pi/7) * DIR
    const R s2 = (R)DIR * (R)0.97492791218182360701813168299393;  // sin(4pi/7) * DIR
    const R s3 = (R)DIR * (R)0.43388373911755812047576833284836;  // sin(6pi/7) * DIR
    const cpx<R> a = u[0];
    const cpx<R> t1 = c_add(u[1], u[6]), t2 = c_add(u[2], u[5]), t3 = c_add(u[3], u[4]);
    const cpx<R> d1 = c_sub(u[1], u[6]), d2 = c_sub(u[2], u[5]), d3 = c_sub(u[3], u[4]);
    u[0] = c_add(a, c_add(t1, c_add(t2, t3)));
    // cos(2 pi jk/7) and sin(2 pi jk/7) for jk mod 7 = 1..6: cos(7-m) = cos(m), sin(7-m) = -sin(m)
    const cpx<R> a1 = {a.x + c1 * t1.x + c2 * t2.x + c3 * t3.x, a.y + c1 * t1.y + c2 * t2.y + c3 * t3.y};
    const cpx<R> a2 = {a.x + c2 * t1.x + c3 * t2.x + c1 * t3.x, a.y + c2 * t1.y + c3 * t2.y + c1 * t3.y};
    const cpx<R> a3 = {a.x + c3 * t1.x + c1 * t2.x + c2 * t3.x, a.y + c3 * t1.y + c1 * t2.y + c2 * t3.y};
    // DIR-signed b_k: k = 1: (s1, s2, s3); k = 2: (s2, -s3, -s1); k = 3: (s3, -s1, s2)
    const cpx<R> b1 = {s1 * d1.x + s2 * d2.x + s3 * d3.x, s1 * d1.y + s2 * d2.y + s3 * d3.y};
    const cpx<R> b2 = {s2 * d1.x - s3 * d2.x - s1 * d3.x, s2 * d1.y - s3 * d2.y - s1 * d3.y};
    const cpx<R> b3 = {s3 * d1.x - s1 * d2.x + s2 * d3.x, s3 * d1.y - s1 * d2.y + s2 * d3.y};
    // i * b = (-b.y, b.x)
    u[1] = {a1.x - b1.y, a1.y + b1.x};
    u[6] = {a1.x + b1.y, a1.y - b1.x};
    u[2] = {a2.x - b2.y, a2.y + b2.x};
    u[5] = {a2.x + b2.y, a2.y - b2.x};
    u[3] = {a3.x - b3.y, a3.y + b3.x};
    u[4] = {a3.x + b3.y, a3.y - b3.x};
}

template <int RAD, int DIR, typename R>
__device__ __forceinline__ void dft(cpx<R> *u)
{
    if constexpr (RAD == 2) dft2<DIR>(u);
    else if constexpr (RAD == 3) dft3<DIR>(u);
    else if constexpr (RAD == 4) dft4<DIR>(u);
    else if constexpr (RAD == 5) dft5<DIR>(u);
    else if constexpr (RAD == 7) dft7<DIR>(u);
    else if constexpr (RAD == 8) dft8<DIR>(u);
    else if constexpr (RAD == 16) dft16<DIR>(u);
}

// ---------------------------------------------------------------------------------------
// LDS layout: one padding slot every 16 complex elements, so the stride-R writes of a
// radix-R pass spread over all banks.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ constexpr int padc(int i) { return i + (i >> 4); }

template <int M>
constexpr int line_slots() { return M + M / 16 + 1; }

// Twiddle seeds of the passes after the first: one base twiddle W_M^(k*S) per butterfly,
// loaded from the FP64-computed table once at kernel start; powers follow by products.
template <int M, int T, int IDX = 0, int P = 1>
constexpr int count_tw()
{
    constexpr RadixPlan plan = make_plan(M, M / T);
    if constexpr (IDX >= plan.n) return 0;
    else {
        constexpr int RAD = plan.r[IDX];
        constexpr int NB = (M / T) / RAD;
        return (P > 1 ? NB : 0) + count_tw<M, T, IDX + 1, P * RAD>();
    }
}

template <typename R, int M, int T, int IDX, int P, int OFF>
__device__ __forceinline__ void load_tw_from(cpx<R> *w, int t, const double *__restrict__ tw)
{
    constexpr RadixPlan plan = make_plan(M, M / T);
    if constexpr (IDX < plan.n) {
        constexpr int RAD = plan.r[IDX];
        constexpr int NB = (M / T) / RAD;
        if constexpr (P > 1) {
            constexpr int S = M / (P * RAD);
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const int i = t + T * q;
                const int e = ((i % P) * S) % M;
                w[OFF + q] = {(R)tw[2 * e], (R)tw[2 * e + 1]};
            }
            load_tw_from<R, M, T, IDX + 1, P * RAD, OFF + NB>(w, t, tw);
        } else {
            load_tw_from<R, M, T, IDX + 1, P * RAD, OFF>(w, t, tw);
        }
    }
}

template <typename R, int M, int T>
struct FFTSeeds {
    static constexpr int NTW = count_tw<M, T>();
    cpx<R> w[NTW > 0 ? NTW : 1];
    // tw: FP64 table exp(-2 pi i e / M), interleaved
    __device__ __forceinline__ void load(int t, const double *__restrict__ tw)
    {
        if constexpr (NTW > 0) load_tw_from<R, M, T, 0, 1, 0>(w, t, tw);
    }
};

// ---------------------------------------------------------------------------------------
// One Stockham pass.  v[E]: this thread's elements t + T*m of the pass input.
// p = product of the radices already applied.  w1[q]: base twiddle of butterfly q.
// LAST: results stay in v (same layout).  Otherwise results are written to `lds`
// (the line's padded complex slice) and read back in the t + T*m layout.
// SYNC: block barriers around the exchange (false when the block is a single wave).
// ---------------------------------------------------------------------------------------
template <typename R, int M, int T, int RAD, int P, bool LAST, int DIR, bool SYNC>
__device__ __forceinline__ void stockham_pass(cpx<R> *v, cpx<R> *lds, int t, const cpx<R> *w1)
{
    constexpr int E = M / T;
    constexpr int NB = E / RAD;  // butterflies per thread
    static_assert(E % RAD == 0, "radix must divide elements per thread");
    cpx<R> out[LAST ? E : 1];
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int i = t + T * q;
        const int k = (P == 1) ? 0 : (i % P);
        cpx<R> u[RAD];
#pragma unroll
        for (int r = 0; r < RAD; r++) u[r] = v[q + r * NB];
        if constexpr (P > 1) {
            cpx<R> wp[RAD];
            wp[1] = (DIR < 0) ? w1[q] : c_conj(w1[q]);
#pragma unroll
            for (int r = 2; r < RAD; r++) wp[r] = (r % 2 == 0) ? c_mul(wp[r / 2], wp[r / 2]) : c_mul(wp[r - 1], wp[1]);
#pragma unroll
            for (int r = 1; r < RAD; r++) u[r] = c_mul(u[r], wp[r]);
        }
        dft<RAD, DIR>(u);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < RAD; r++) out[q + r * NB] = u[r];
        } else {
            const int j = (i / P) * P * RAD + k;
#pragma unroll
            for (int r = 0; r < RAD; r++) lds[padc(j + r * P)] = u[r];
        }
    }
    if constexpr (LAST) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = out[m];
    } else {
        if constexpr (SYNC) __syncthreads();
        else __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = lds[padc(t + T * m)];
        if constexpr (SYNC) __syncthreads();
        else __builtin_amdgcn_wave_barrier();
    }
}

template <typename R, int M, int T, int DIR, bool SYNC, int IDX, int P, int OFF>
__device__ __forceinline__ void stockham_from(cpx<R> *v, cpx<R> *lds, int t, const cpx<R> *w)
{
    constexpr RadixPlan plan = make_plan(M, M / T);
    if constexpr (IDX < plan.n) {
        constexpr int RAD = plan.r[IDX];
        constexpr int NB = (M / T) / RAD;
        stockham_pass<R, M, T, RAD, P, IDX == plan.n - 1, DIR, SYNC>(v, lds, t, w + OFF);
        stockham_from<R, M, T, DIR, SYNC, IDX + 1, P * RAD, OFF + (P > 1 ? NB : 0)>(v, lds, t, w);
    }
}

// Full FFT of one line: v in/out in the t + T*m layout.  With SYNC all threads of the
// block must call this together (block barriers around each exchange).
template <typename R, int M, int T, int DIR, bool SYNC = true>
__device__ __forceinline__ void line_fft(cpx<R> *v, cpx<R> *lds, int t, const FFTSeeds<R, M, T> &s)
{
    static_assert(plan_ok(M, M / T), "no radix plan for this (M, T)");
    stockham_from<R, M, T, DIR, SYNC, 0, 1, 0>(v, lds, t, s.w);
}

}  // namespace pp
