// pp_line_bs.h -- line transforms of any length N (2 <= N <= 1024) for CDNA4 (gfx950):
// lengths the compiled Stockham plans of pp_line.h do not cover (odd N, N/2 with a prime
// factor above 7, N/2 not in the compiled list).
//
// The reference gets every length from FFTW / cuFFT plans of any n (src/poissonCPU.cpp:56,87,
// 162,211, src/poissonGPU.cpp:13,37) and SignalProc handles odd sizes
// (src/signalProc.cpp:34-44,161-195).  Here every kind is written as one chirp sum over the
// line,
//     S_k = sum_{j<N} a_j exp(-+ i pi beta j k),   beta = 1/N (DCT/DST kinds) or 2/N (PP),
// between a pre-twiddle (a_j from the real input) and a post-twiddle (real output from S_k),
// and the chirp sum is Bluestein's convolution: with c_m = exp(+- i pi beta m^2 / 2),
//     S_k = conj(c_k) sum_j (a_j conj(c_j)) c_{k-j},
// a circular convolution of length L >= 2N - 1 (L a power of two) done as forward FFT,
// product with the precomputed FFT of the chirp (divided by L), inverse FFT -- both through
// the Stockham engine of pp_fft.h.  The host folds the pre-/post-twiddles and the chirp into
// FP64 tables per family and direction (P_j, B_j, Q_k; pp_setup.h fill_bs_tables); the
// families are DCT/DST-II/III (NN, DD), DCT/DST-IV (ND, DN) and the real DFT (PP):
//
//   kind / direction      input a_j                     output (position)
//   NN forward  DCT-II    x_j P_j                        Re(Q_k S'_k)          (k)
//   DD forward  DST-II    (-1)^j x_j P_j                 Re(Q_k S'_k)          (N-1-k)
//   NN inverse  DCT-III/2 x_j P_j                        Re(Q_k S'_k)          (k)
//   DD inverse  DST-III/2 x_{N-1-j} P_j                  (-1)^k Re(Q_k S'_k)   (k)
//   ND          DCT-IV    x_j P_j                        Re(Q_k S'_k)          (k)
//   DN          DST-IV    x_j P_j                        -Im(Q_k S'_k)         (k)
//   PP forward  R2HC      x_j P_j                        Re / -Im(Q_k S'_k)    (k <= N/2 / k > N/2)
//   PP inverse  HC2R      X_j P_j (X from half-complex)  Re(Q_k S'_k)          (k)
//
// where S' is the raw convolution.  (DST-II via DCT-II of (-1)^j x_j, reversed, and
// DST-IV_k = -Im of the DCT-IV sum, are the identities SignalProc::swap relies on,
// src/signalProc.cpp:235-265.)  Scalings match pp_line.h (forward DCT-II / DST-II /
// DCT-IV / DST-IV x2, inverse /2 for DCT-III / DST-III, round trip N per axis).  Launch
// arguments, fused scalings, the Dirichlet boundary term, line sums, accumulation and the
// z-tiled y layout are those of k_line (LineArgs, line_block, line_addr).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pp_fft.h"
#include "pp_line.h"

namespace pp {

// convolution length for line length N: the smallest power of two >= max(16, 2N - 1)
constexpr int bs_len(int N)
{
    int L = 16;
    while (L < 2 * N - 1) L *= 2;
    return L;
}
constexpr int kBsMaxN = 1024;
constexpr int bs_threads(int L) { return L / 8; }  // 8 complex elements per thread

// columns per block of a y launch / lines per block of an x launch.  The chirp convolution
// computes in FP64 in every precision form (R is the storage type of the FP32 forms and does
// not change the shape): its two length-L FFTs and kernel product carry FP32 rounding well
// above the compiled plans', more than two refinement steps of the mixed form remove, while
// FP32 arithmetic issues at the FP64 rate here (packed FP32 is off), so FP64 costs only LDS
// and registers.
template <typename R, int L, int AX>
constexpr int bs_lines()
{
    constexpr int T = bs_threads(L);
    constexpr int bytes = line_slots<L>() * 2 * (int)sizeof(double);
    int lines = AX == 0 ? 256 / T : 16;
    if (lines < 1) lines = 1;
    while (lines > 1 && lines * bytes > 144 * 1024) lines /= 2;
    // blocks of at most 512 threads (up to 256 VGPRs: the long lines need ~140)
    while (lines > 1 && lines * T > 512) lines /= 2;
    return lines;
}

template <typename TIn, typename TOut, typename R, int L, int LINES, int AX, int FWD, int KIND>
__global__ void __launch_bounds__(bs_threads(L) * LINES) k_line_bs(LineArgs a0)
{
    constexpr int T = bs_threads(L);
    constexpr int E = L / T;
    constexpr int H = E / 2;  // N <= L / 2 = T * H: elements t + T c, c < H, cover the line
    constexpr int LS = line_slots<L>();
    LineArgs a = a0;
    int bx, by;
    line_block<TIn, TOut, R, LINES, AX>(a0, a, bx, by);
    const int N = a0.n;
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
    cpx<R> *Ls = lds + l * LS;

    const TIn *in = reinterpret_cast<const TIn *>(a.in);
    TOut *out = reinterpret_cast<TOut *>(a.out);
    const LineAddr ad = line_addr<LINES, AX>(a, N, bx, by, l);

    // table sections of this family and direction (host layout, pp_setup.h fill_bs_tables:
    // P and Q per family {II/III, IV, PP} x {forward, inverse}, N complex each; B per chirp
    // {half-angle, full} x {forward, inverse}, L complex each)
    constexpr int FAM = (KIND == KIND_NN || KIND == KIND_DD) ? 0 : (KIND == KIND_PP ? 2 : 1);
    constexpr int SEC = 2 * FAM + (FWD ? 0 : 1);
    const double *P = a.A + (long)2 * N * SEC;
    const double *Q = a.C + (long)2 * N * SEC;
    // (DCT/DST-IV is its own inverse: both directions use the forward chirp)
    constexpr int BSEC = FAM == 2 ? 2 + (FWD ? 0 : 1) : (FAM == 0 ? (FWD ? 0 : 1) : 0);
    const double *B = a.B + (long)2 * L * BSEC;

    // ---- loads: element n = t + T c ----
    double xin[H];
    {
        const bool tin = (AX == 1) && (a.ztiled & ZT_IN);
        const long ib = tin ? ad.tbase : ad.base, is = tin ? ad.tstride : ad.stride;
#pragma unroll
        for (int c = 0; c < H; c++) {
            const int n = t + T * c;
            const double v = (double)in[ib + (long)(n < N ? n : 0) * is];
            xin[c] = n < N ? v : 0.0;
        }
    }
    // x inverse, FP32 -> FP64: the solution being accumulated into and the plane term are
    // read with the input
    constexpr bool XINV = (AX == 0) && !FWD && !std::is_same<TIn, TOut>::value;
    double accv[XINV ? H : 1];
    double padd = 0.0;
    if constexpr (XINV) {
        if (a.accumulate) {
#pragma unroll
            for (int c = 0; c < H; c++) {
                const int n = t + T * c;
                accv[c] = n < N ? (double)out[ad.base + (long)n * ad.stride] : 0.0;
            }
        } else if (a.acc0) {
#pragma unroll
            for (int c = 0; c < H; c++) {
                const int n = t + T * c;
                const long i = ad.base + (long)(n < N ? n : 0) * ad.stride;
                accv[c] = n < N ? (double)a.acc0[i] + (a.acc1 ? (double)a.acc1[i] : 0.0) : 0.0;
            }
        }
        if (a.plane_add) padd = a.plane_add[ad.lz];
    }
    // convolution kernel spectrum of this thread's elements (independent of the data)
    cpx<R> bk[E];
#pragma unroll
    for (int m = 0; m < E; m++) bk[m] = to_c<R>(ld64(B, t + T * m));
    FFTSeeds<R, L, T> fs;
    fs.load(t, a.tw);

    // ---- fused input scaling / Dirichlet boundary term / line sums ----
    if constexpr (AX == 0 && FWD) {
        const DirLine dl = dirichlet_line(a, ad.ly, ad.lz);
#pragma unroll
        for (int c = 0; c < H; c++) {
            const int n = t + T * c;
            xin[c] = xin[c] * a.in_scale;  // rhs * c3 * c3 (src/pencilDcmp.cpp:5639)
            xin[c] = xin[c] * a.in_scale;
            if (dl.cyz != 0.0 && n < N) xin[c] = xin[c] - gld(a.dtab, n) * dl.cyz;
            if (n == 0) xin[c] = xin[c] - dl.cx0;
            if (n == N - 1) xin[c] = xin[c] - dl.cxn;
        }
        if (a.line_sums) {
            double lsum = 0.0;
#pragma unroll
            for (int c = 0; c < H; c++) lsum += xin[c];
            if constexpr (T <= 64) {
                // the T threads of a line are adjacent lanes of one wave
#pragma unroll
                for (int off = T / 2; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, T);
            } else {
                // a line spans several waves: through its LDS slice (free until staging)
                double *red = reinterpret_cast<double *>(Ls);
                red[t] = lsum;
                __syncthreads();
                if (t == 0)
                    for (int i = 1; i < T; i++) lsum += red[i];
                __syncthreads();
            }
            if (t == 0 && ad.valid) a.line_sums[(long)bx * LINES + l] = lsum;
        }
    } else {
#pragma unroll
        for (int c = 0; c < H; c++) xin[c] = xin[c] * a.in_scale;
    }

    // ---- stage the real line in LDS (the inputs of a_j are permuted / paired) ----
#pragma unroll
    for (int c = 0; c < H; c++) {
        const int n = t + T * c;
        if (n < N) lr(Ls, n) = (R)xin[c];
    }
    __syncthreads();

    // ---- a_j = (input) * P_j, zero from N to L ----
    cpx<R> v[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
        const int j = t + T * m;
        v[m] = cpx<R>{(R)0, (R)0};
        if (m < H && j < N) {
            const cpx<R> p = to_c<R>(ld64(P, j));
            if constexpr (KIND == KIND_PP && !FWD) {
                // X_j from the half-complex input: Re at j, Im at N - j (conjugate above N/2)
                cpx<R> X;
                if (2 * j < N) X = {lr(Ls, j), j == 0 ? (R)0 : lr(Ls, N - j)};
                else if (2 * j == N) X = {lr(Ls, j), (R)0};
                else X = {lr(Ls, N - j), -lr(Ls, j)};
                v[m] = c_mul(X, p);
            } else {
                R x;
                if constexpr (KIND == KIND_DD && !FWD) x = lr(Ls, N - 1 - j);
                else x = lr(Ls, j);
                if constexpr (KIND == KIND_DD && FWD) x = (j & 1) ? -x : x;
                v[m] = {x * p.x, x * p.y};
            }
        }
    }
    __syncthreads();

    // ---- circular convolution with the chirp: FFT, spectrum product, inverse FFT ----
    line_fft<R, L, T, -1, true>(v, Ls, t, fs);
#pragma unroll
    for (int m = 0; m < E; m++) v[m] = c_mul(v[m], bk[m]);
    line_fft<R, L, T, +1, true>(v, Ls, t, fs);

    // ---- post-twiddle, real output, fused output scaling ----
    if (!ad.valid) return;
#pragma unroll
    for (int m = 0; m < H; m++) {
        const int k = t + T * m;
        if (k >= N) continue;
        const cpx<R> y = c_mul(to_c<R>(ld64(Q, k)), v[m]);
        double val;
        if constexpr (KIND == KIND_DN) val = -(double)y.y;
        else if constexpr (KIND == KIND_PP && FWD) val = (2 * k <= N) ? (double)y.x : -(double)y.y;
        else if constexpr (KIND == KIND_DD && !FWD) val = (k & 1) ? -(double)y.x : (double)y.x;
        else val = (double)y.x;
        const int pos = (KIND == KIND_DD && FWD) ? N - 1 - k : k;
        val = a.out_scale * val;
        if constexpr (AX == 0 && !FWD) {
            if constexpr (XINV) {
                if (a.plane_add) val += padd;
                if (a.accumulate || a.acc0) val += accv[m];
            } else {
                if (a.plane_add) val += a.plane_add[ad.lz];
                if (a.accumulate) val += (double)out[ad.base + (long)pos * ad.stride];
            }
        }
        if (AX == 1 && (a.ztiled & ZT_OUT)) out[ad.tbase + (long)pos * ad.tstride] = (TOut)val;
        else out[ad.base + (long)pos * ad.stride] = (TOut)val;
    }
}

template <typename TIn, typename TOut, typename R, int L, int AX, int FWD, int KIND>
hipError_t launch_line_bs_cfg(const LineArgs &a, hipStream_t s)
{
    constexpr int T = bs_threads(L);
    constexpr int LINES = bs_lines<R, L, AX>();
    const size_t lds = (size_t)LINES * line_slots<L>() * sizeof(cpx<R>);
    auto kern = k_line_bs<TIn, TOut, R, L, LINES, AX, FWD, KIND>;
    if (lds > 64 * 1024) {
        static bool attr_set = false;
        if (!attr_set) {
            hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr_set = true;
        }
    }
    dim3 grid;
    if (a.items) grid = dim3((unsigned)a.nlines);  // batched: nlines holds the item count
    else if (AX == 0) grid = dim3((unsigned)((a.nlines + LINES - 1) / LINES));
    else grid = dim3((unsigned)((a.nlines + LINES - 1) / LINES), (unsigned)a.nplanes);
    hipLaunchKernelGGL(kern, grid, dim3(T * LINES), lds, s, a);
    return hipGetLastError();
}

#define PP_BS_LENGTHS(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

template <typename TIn, typename TOut, typename R, int AX, int FWD, int KIND>
hipError_t launch_line_bs_sz(const LineArgs &a, hipStream_t s)
{
    if (a.n < 2 || a.n > kBsMaxN) return hipErrorInvalidValue;
    switch (bs_len(a.n)) {
#define PPL(l) \
    case l: return launch_line_bs_cfg<TIn, TOut, R, l, AX, FWD, KIND>(a, s);
        PP_BS_LENGTHS(PPL)
#undef PPL
    default: return hipErrorInvalidValue;
    }
}

// every precision form, axis and direction of one kind (LineForm as in launch_line_kind)
template <int KIND>
hipError_t launch_line_bs_kind(int form, int AX, int FWD, const LineArgs &a, hipStream_t s)
{
    switch (form) {
    case FORM_F64:
        if (AX == 0) return FWD ? launch_line_bs_sz<double, double, double, 0, 1, KIND>(a, s)
                                : launch_line_bs_sz<double, double, double, 0, 0, KIND>(a, s);
        return FWD ? launch_line_bs_sz<double, double, double, 1, 1, KIND>(a, s)
                   : launch_line_bs_sz<double, double, double, 1, 0, KIND>(a, s);
    // FP32 storage, FP64 chirp convolution (bs_lines)
    case FORM_F32:
        if (AX == 0) return FWD ? launch_line_bs_sz<float, float, double, 0, 1, KIND>(a, s)
                                : launch_line_bs_sz<float, float, double, 0, 0, KIND>(a, s);
        return FWD ? launch_line_bs_sz<float, float, double, 1, 1, KIND>(a, s)
                   : launch_line_bs_sz<float, float, double, 1, 0, KIND>(a, s);
    case FORM_F64_TO_F32:
        if (AX == 0 && FWD) return launch_line_bs_sz<double, float, double, 0, 1, KIND>(a, s);
        return hipErrorInvalidValue;
    case FORM_F32_TO_F64:
        if (AX == 0 && !FWD) return launch_line_bs_sz<float, double, double, 0, 0, KIND>(a, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

// lines (x) or columns (y) per block of the arbitrary-length launch for line length N
template <typename R>
inline int bs_block_lines(int N, int AX)
{
    if (N < 2 || N > kBsMaxN) return 0;
    switch (bs_len(N)) {
#define PPL(l) \
    case l: return AX == 0 ? bs_lines<R, l, 0>() : bs_lines<R, l, 1>();
        PP_BS_LENGTHS(PPL)
#undef PPL
    default: return 0;
    }
}

// every line length 2..1024: the compiled plans of pp_line.h, the rest through this file
inline bool line_size_supported(int n) { return line_compiled(n) || (n >= 2 && n <= kBsMaxN); }

// lines (x) or columns (y) per block of whichever launch serves length N
template <typename R>
inline int line_lines_any(int N, int AX)
{
    return line_compiled(N) ? line_block_lines<R>(N, AX) : bs_block_lines<R>(N, AX);
}

}  // namespace pp
