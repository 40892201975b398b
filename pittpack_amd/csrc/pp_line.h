// pp_line.h -- batched real-to-real line transforms along x (contiguous lines) or y
// (strided columns) for CDNA4 (gfx950).
//
// Replaces the reference's per-axis transform step: SignalProc's Makhoul pre/post
// processing (src/signalProc.cpp:32-265) around cuFFT Z2Z (src/poissonGPU.cpp:5-98), the
// x<->y plane transposes (src/pencilDcmp.cpp:2279-2515), assignRhs's dz^2 scaling
// (5591-5647), modifyRhsDirichlet (4623-4835) and rescale (3503-3531).
//
// Every axis keeps a real line of length N = 2M real:
//   PP : R2HC (half-complex)          / HC2R          (unnormalised, round trip N)
//   NN : DCT-II  (2 sum x cos ...)    / DCT-III / 2   (round trip N)
//   DD : DST-II                       / DST-III / 2
//   ND : DCT-IV  (x2)                 / DCT-IV        (quarter-wave, round trip N)
//   DN : DST-IV  (x2)                 / DST-IV
// each through one complex FFT of length M on the packed line (z_m = v_{2m} + i v_{2m+1}).
//
// Pipeline inside one block: coalesced global load (+ fused input scaling / Dirichlet
// term) -> scatter into the padded LDS slice of the line -> pre-processing into registers
// -> Stockham passes -> post-processing -> coalesced global store (+ fused output scale).
// Twiddles come from FP64 tables: a handful of seeds per thread are loaded at kernel
// start and the per-element values follow by FP64 products, so no table load sits between
// two barriers.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pp_fft.h"
#include "pp_xcd.h"

namespace pp {

enum AxisKind : int { KIND_DD = 0, KIND_NN = 1, KIND_PP = 2, KIND_ND = 3, KIND_DN = 4 };

// Batched launches: one launch covers the blocks of many independent solves of the same
// line length and kind.  Block b works on items[b] of solve `solve`, whose fields sit at
// `offset` in the launch's in/out arenas.
struct LineSolve {
    long offset;       // element offset of the solve's field in the in/out arenas
    int nx, ny, nz;
    int dmask;         // Dirichlet faces (x forward with BF_RHS)
    double dz;         // x forward with BF_RHS: input * dz * dz
    double inv_scale;  // x inverse with BF_SCALE: 1/(nx*ny)
    const double *dtab;
    long sums_off;   // offset of the solve's ny*nz line sums in line_sums
    long plane_off;  // offset of the solve's nz (+2) plane values in plane_add
    long toffset;    // z-tiled launches: offset of the solve's z tiles on the tiled side
    int tnz;         // rows per z tile; 0: the solve has no z tiles (its field stays in place)
};
struct LineItem {
    int solve, bx, by;
};
enum : int { BF_RHS = 1, BF_SCALE = 2 };

struct LineArgs {
    const void *in;
    void *out;
    long nlines;  // x: number of lines (ny*nz); y: columns per plane (nx)
    int ncols;    // y: element stride (= nx)
    int nplanes;  // y: nz
    // FP64 tables for this axis, interleaved complex:
    //   tw[M] = exp(-2 pi i e/M), A[M+1] = exp(-2 pi i k/N), B[M+1] = exp(-i pi k/(2N)),
    //   C[M] = exp(-i pi (4n+1)/(4N)), D[M] = exp(-i pi k/N)
    const double *tw, *A, *B, *C, *D;
    double in_scale;   // forward x: dz (applied twice: rhs*dz*dz); otherwise 1
    double out_scale;  // inverse x: 1/(Nx*Ny); otherwise 1
    // Dirichlet boundary term (forward x only): dtab = [ex(nx) ey(ny) ez(nz) exf0 exf1 eyf0
    // eyf1 ezf0 ezf1], dmask bit f set when face f (xmin,xmax,ymin,ymax,zmin,zmax) is D.
    const double *dtab;
    int dmask;
    int nx, ny, nz;
    // x forward only: FP64 sum over each line of the (scaled, Dirichlet-corrected) input,
    // line_sums[y + ny*z] (null: not wanted)
    double *line_sums;
    // x inverse only: out = (accumulate ? out : 0) + out_scale * value + plane_add[z]
    const double *plane_add;
    int accumulate;
    // x inverse from FP32 to FP64 (the mixed form's last store): out = acc0 + acc1 +
    // out_scale * value + plane_add[z], the FP32 parts of the solution summed in FP64 (acc1
    // may be null; null acc0: no such parts)
    const float *acc0, *acc1;
    // batched launch (null: one solve, grid = blocks): per-block items, per-solve fields,
    // and which per-solve scalings apply (BF_RHS, BF_SCALE)
    const LineItem *items;
    const LineSolve *solves;
    int bflags;
    // y only: z-tiled layout of the output (ZT_OUT, forward) or input (ZT_IN, inverse):
    // element (x, y, z) at ((x/W * ny + y) * tnz + z) * W + x%W with W = tile_w (else columns
    // per block) and tnz >= nz rows per tile, so each (x/W, y) tile is contiguous for the z solve
    int ztiled;
    int tnz;
    int tile_w;  // tile width W (0: LINES, the block's column count); a multiple of LINES
    // host side, FP64 form: 1 = the alternative launch shape (line_threads_alt), 0 = default
    int variant;
    // line length; read by the arbitrary-length kernels (pp_line_bs.h), whose length is not a
    // template parameter
    int n;
};
enum : int { ZT_OUT = 1, ZT_IN = 2 };

struct c64 {
    double x, y;
};
__device__ __forceinline__ c64 m64(c64 a, c64 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c64 ld64(const double *t, int k) { return {t[2 * k], t[2 * k + 1]}; }
template <typename R>
__device__ __forceinline__ cpx<R> to_c(c64 a) { return {(R)a.x, (R)a.y}; }

// load through the global address space (pointers taken from descriptor tables are generic)
__device__ __forceinline__ double gld(const double *p, long i)
{
    return ((const __attribute__((address_space(1))) double *)p)[i];
}

// position in the real view of a line slice where input element n is placed
template <int N, int KIND, int FWD>
__device__ __forceinline__ int load_pos(int n, bool &neg)
{
    neg = false;
    if constexpr (KIND == KIND_ND || KIND == KIND_DN) {
        const int a = (KIND == KIND_DN) ? (N - 1 - n) : n;
        return (a & 1) ? (N - a) : a;
    } else if constexpr (FWD && (KIND == KIND_NN || KIND == KIND_DD)) {
        if constexpr (KIND == KIND_DD) neg = (n & 1);
        return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
    } else {
        return n;
    }
}

// element held at real-view position p: the inverse of load_pos for forward transforms
// and of store_pos for inverse ones (NN / DD share the Makhoul order both ways)
template <int N, int KIND>
__device__ __forceinline__ int pos_elem(int p, bool &neg)
{
    neg = false;
    if constexpr (KIND == KIND_NN || KIND == KIND_DD) {
        const int n = (p < N / 2) ? 2 * p : 2 * (N - 1 - p) + 1;
        if constexpr (KIND == KIND_DD) neg = (n & 1);
        return n;
    } else if constexpr (KIND == KIND_ND) {
        return (p & 1) ? N - p : p;
    } else if constexpr (KIND == KIND_DN) {
        return (p & 1) ? p - 1 : N - 1 - p;
    } else {
        return p;
    }
}

template <int N, int KIND, int FWD>
__device__ __forceinline__ int store_pos(int n, bool &neg)
{
    neg = false;
    if constexpr (!FWD && (KIND == KIND_NN || KIND == KIND_DD)) {
        if constexpr (KIND == KIND_DD) neg = (n & 1);
        return (n & 1) ? (N - 1 - (n >> 1)) : (n >> 1);
    } else {
        return n;
    }
}

template <typename R>
__device__ __forceinline__ R &lr(cpx<R> *L, int p)
{
    return reinterpret_cast<R *>(&L[padc(p >> 1)])[p & 1];
}

// Block coordinates of a line launch.  y blocks narrower than a cache line keep the tiles of
// one plane on one XCD; a batched launch (a0.items) takes its block's item and the item's
// solve: per-solve fields, sizes and scalings are resolved into `a`.
template <typename TIn, typename TOut, typename R, int LINES, int AX>
__device__ __forceinline__ void line_block(const LineArgs &a0, LineArgs &a, int &bx, int &by)
{
    bx = blockIdx.x;
    by = blockIdx.y;
    if constexpr (AX == 1 && LINES * sizeof(R) < 128) {
        if (!a0.items) {
            const int lin = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
            bx = lin % gridDim.x;
            by = lin / gridDim.x;
        }
    }
    if (a0.items) {
        const LineItem it = a0.items[blockIdx.x];
        const LineSolve sv = a0.solves[it.solve];
        if (a0.ztiled && sv.tnz == 0) {
            // a solve without z tiles (periodic z) keeps its field in place: the y transform
            // reads and writes the launch's other arena (tiled launches are FP64 -> FP64)
            TOut *f = (a0.ztiled & ZT_OUT) ? reinterpret_cast<TOut *>(const_cast<void *>(a0.in))
                                           : reinterpret_cast<TOut *>(a0.out);
            a.in = f + sv.offset;
            a.out = f + sv.offset;
            a.ztiled = 0;
        } else {
            a.in = reinterpret_cast<const TIn *>(a0.in) + ((a0.ztiled & ZT_IN) ? sv.toffset : sv.offset);
            a.out = reinterpret_cast<TOut *>(a0.out) + ((a0.ztiled & ZT_OUT) ? sv.toffset : sv.offset);
        }
        a.tnz = sv.tnz;
        a.nx = sv.nx;
        a.ny = sv.ny;
        a.nz = sv.nz;
        if constexpr (AX == 0) {
            a.nlines = (long)sv.ny * sv.nz;
        } else {
            a.nlines = sv.nx;
            a.ncols = sv.nx;
            a.nplanes = sv.nz;
        }
        a.in_scale = (a0.bflags & BF_RHS) ? sv.dz : 1.0;
        a.dmask = (a0.bflags & BF_RHS) ? sv.dmask : 0;
        a.dtab = sv.dtab;
        a.out_scale = (a0.bflags & BF_SCALE) ? sv.inv_scale : 1.0;
        if (a0.line_sums) a.line_sums = a0.line_sums + sv.sums_off;
        if (a0.plane_add) a.plane_add = a0.plane_add + sv.plane_off;
        if (a0.acc0) a.acc0 = a0.acc0 + sv.offset;
        if (a0.acc1) a.acc1 = a0.acc1 + sv.offset;
        bx = it.bx;
        by = it.by;
    }
}

// Where line l of block (bx, by) lives: element n at base + n * stride (tbase + n * tstride
// on the z-tiled side of a y transform).  Lines past the end of the grid map to line 0 (in
// bounds) with valid = false: they are transformed like the others and never stored.
struct LineAddr {
    long base, stride, tbase, tstride;
    bool valid;
    int ly, lz;  // x lines: the line's y and z
    // y columns: base = ubase + ucol and tbase = tubase + tcol, split into the block's uniform
    // offset and the lane's (small, non-negative) column offset
    long ubase, tubase;
    int ucol, tcol;
};

template <int LINES, int AX>
__device__ __forceinline__ LineAddr line_addr(const LineArgs &a, int N, int bx, int by, int l)
{
    LineAddr d{};
    if constexpr (AX == 0) {
        long line = (long)bx * LINES + l;
        d.valid = line < a.nlines;
        if (!d.valid) line = 0;
        d.base = line * N;
        d.stride = 1;
        d.ly = (int)(line % a.ny);
        d.lz = (int)(line / a.ny);
    } else {
        // columns past the grid read the last column (in bounds)
        const int col0 = bx * LINES;
        d.valid = col0 + l < a.ncols;
        d.ubase = (long)by * N * a.ncols + col0;
        d.ucol = d.valid ? l : a.ncols - 1 - col0;
        d.base = d.ubase + d.ucol;
        d.stride = a.ncols;
        if (a.ztiled) {
            if (a.tile_w == 0 || a.tile_w == LINES) {
                // tile (bx, n) holds rows z of LINES columns: ((bx * ny + n) * tnz + z) * LINES + l
                d.tubase = ((long)bx * N * a.tnz + by) * LINES;
                d.tstride = (long)a.tnz * LINES;
            } else {
                // tiles W = tile_w columns wide (a multiple of LINES): this block fills LINES of
                // tile bx / (W / LINES); neighbouring blocks, which complete each row, are
                // numbered onto one XCD (W * 8 bytes is one 128-byte line)
                const int W = a.tile_w, sub = W / LINES;
                d.tubase = ((long)(bx / sub) * N * a.tnz + by) * W + (bx % sub) * LINES;
                d.tstride = (long)a.tnz * W;
            }
            d.tcol = l;
            d.tbase = d.tubase + d.tcol;
        }
    }
    return d;
}

// Dirichlet boundary term of one x line (modifyRhsDirichlet, src/pencilDcmp.cpp:4623-4835):
// the z- and y-face terms are ex[n] times a line constant cyz and the x-face terms hit
// elements 0 (cx0) and N-1 (cxn) only, so the line constants are formed once per thread
struct DirLine {
    double cyz = 0.0, cx0 = 0.0, cxn = 0.0;
};

__device__ __forceinline__ DirLine dirichlet_line(const LineArgs &a, int ly, int lz)
{
    DirLine d;
    if (a.dmask) {
        const double *ey = a.dtab + a.nx, *ez = a.dtab + a.nx + a.ny;
        const double *ef = a.dtab + a.nx + a.ny + a.nz;
        const double eyl = gld(ey, ly), ezl = gld(ez, lz);
        if ((a.dmask & 16) && lz == 0) d.cyz += 2. * eyl * gld(ef, 4);
        if ((a.dmask & 32) && lz == a.nz - 1) d.cyz += 2. * eyl * gld(ef, 5);
        if ((a.dmask & 4) && ly == 0) d.cyz += 2. * gld(ef, 2) * ezl;
        if ((a.dmask & 8) && ly == a.ny - 1) d.cyz += 2. * gld(ef, 3) * ezl;
        if (a.dmask & 1) d.cx0 = 2. * gld(ef, 0) * eyl * ezl;
        if (a.dmask & 2) d.cxn = 2. * gld(ef, 1) * eyl * ezl;
    }
    return d;
}

// Inputs are held in their own precision (FP32 inputs are only ever scaled by 1): half the
// registers for the FP32 forms
template <typename TIn>
using line_xt = typename std::conditional<std::is_same<TIn, float>::value, float, double>::type;

// y forward and y quarter-wave (either direction): lanes are columns, so every row is a
// coalesced load and each thread reads the rows of its own FFT inputs directly (real view
// position p = 2 (t + T m) + c % 2), no LDS scatter; otherwise element n = t + T c,
// scattered through LDS after the load
template <int AX, int FWD, int KIND>
constexpr bool line_direct_in()
{
    return (AX == 1) && (FWD || KIND == KIND_ND || KIND == KIND_DN);
}

// Minimum waves per SIMD a line kernel is compiled for (its register bound).  Power-of-two
// lengths from 512 (8 elements per thread): at least 4 (<= 128 VGPRs), so two 512-thread
// blocks fit a CU; other lengths keep their registers, and so does the FP32 -> FP64 x inverse
// of 1024-point lines (its prefetched solution would spill).  FP32 y blocks of 1024 threads
// whose LDS slice fits a CU twice: 8 (<= 64 VGPRs), so two blocks share a CU and one
// block's loads overlap the other's transform.
template <typename TIn, typename TOut, typename R, int M, int T, int LINES, int AX, int FWD>
constexpr int line_min_waves()
{
    if (AX == 1 && sizeof(R) == 4 && T * LINES == 1024 && 2 * LINES * line_slots<M>() * 8 <= 160 * 1024) return 8;
    return (M / T == 8 && M >= 256 && !(AX == 0 && !FWD && !std::is_same<TIn, TOut>::value && M >= 512)) ? 4 : 1;
}

// x lines whose T threads cover less than a 128-byte line per access (T * sizeof(R) < 128:
// short lines) are staged: the block's lines are one contiguous run of memory, so the block
// stores (loads) it with whole-wave contiguous accesses out of (into) the lines' LDS slices,
// instead of T-element fragments per line and instruction.  Only where the stored precision
// is the LDS precision (the slices hold values of type R); loads of inverse transforms only
// (measured: the staged load slows the forward one, whose fragments feed the fused scaling
// and line sums directly)
template <typename TIn, typename R, int T, int AX, int FWD>
constexpr bool line_stage_in()
{
    return AX == 0 && !FWD && T * (int)sizeof(R) < 128 && sizeof(TIn) == sizeof(R);
}
template <typename TOut, typename R, int T, int AX>
constexpr bool line_stage_out()
{
    return AX == 0 && T * (int)sizeof(R) < 128 && sizeof(TOut) == sizeof(R);
}

// staged load of a block's x lines (line0 .. line0 + LINES - 1, those below nlines): element e
// of the run goes to position e % N of line e / N's slice; then each thread reads its inputs
// n = t + T c from its line's slice (the layout line_load gives)
template <typename TIn, typename R, int M, int T, int LINES>
__device__ __forceinline__ void stage_lines_in(const LineArgs &a, long line0, cpx<R> *lds, int t, int l,
                                               line_xt<TIn> (&xin)[2 * (M / T)])
{
    constexpr int N = 2 * M, LS = line_slots<M>(), NT = T * LINES, C = 2 * (M / T);
    const TIn *src = reinterpret_cast<const TIn *>(a.in) + line0 * N;
    const long rem = a.nlines - line0;
    const int total = (int)(rem < LINES ? rem : LINES) * N;
    // a full block moves N / T = C elements per thread: every load is issued before the
    // first LDS write
    const int tid = (int)threadIdx.x;
    R v[C];
#pragma unroll
    for (int c = 0; c < C; c++) v[c] = (tid + NT * c < total) ? (R)src[tid + NT * c] : (R)0;
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int e = tid + NT * c, l2 = e / N;
        if (e < total) lr(lds + l2 * LS, e - l2 * N) = v[c];
    }
    __syncthreads();
    cpx<R> *L = lds + l * LS;
#pragma unroll
    for (int c = 0; c < 2 * (M / T); c++) xin[c] = (line_xt<TIn>)lr(L, t + T * c);
    __syncthreads();  // the slices are rewritten by the pre-processing next
}

// Issue the coalesced global loads of one line (column) of a block
template <typename TIn, typename R, int M, int T, int AX, int FWD, int KIND>
__device__ __forceinline__ void line_load(const LineArgs &a, const LineAddr &ad, int t, line_xt<TIn> (&xin)[2 * (M / T)])
{
    constexpr int N = 2 * M;
    constexpr int E = M / T;
    const TIn *in = reinterpret_cast<const TIn *>(a.in);
    if constexpr (AX == 1) {
        // element n(c, t) is affine in the thread's row t: the address is a block-uniform
        // pointer (row n(c, t) at its smallest over t, in scalar registers) plus a 32-bit lane
        // byte offset (< 2^28 for lines up to 1024 and strides up to 32768), so no 64-bit
        // address per load is held in vector registers.  The Makhoul order of NN / DD inputs
        // read directly with an odd element count per thread is only piecewise affine (its
        // two halves meet inside a row of threads): those rows take row 0 as the base.
        constexpr bool AFFINE = !(line_direct_in<AX, FWD, KIND>() && (KIND == KIND_NN || KIND == KIND_DD) && E % 2);
        const bool tin = a.ztiled & ZT_IN;
        const long ub = tin ? ad.tubase : ad.ubase, is = tin ? ad.tstride : ad.stride;
        const int lc = tin ? ad.tcol : ad.ucol;
        auto row = [](int c, int tt) {
            int n = tt + T * c;
            if constexpr (line_direct_in<AX, FWD, KIND>()) {
                bool ng;
                n = pos_elem<N, KIND>(2 * (tt + T * (c >> 1)) + (c & 1), ng);
            }
            return n;
        };
#pragma unroll
        for (int c = 0; c < 2 * E; c++) {
            const int nmin = AFFINE ? min(row(c, 0), row(c, T - 1)) : 0;
            const TIn *pu = in + ub + (long)nmin * is;
            const unsigned lane_bytes = (unsigned)(((row(c, t) - nmin) * (int)is + lc) * (int)sizeof(TIn));
            xin[c] = (line_xt<TIn>)*reinterpret_cast<const TIn *>(reinterpret_cast<const char *>(pu) + lane_bytes);
        }
        return;
    }
    const long ib = ad.base, is = ad.stride;
#pragma unroll
    for (int c = 0; c < 2 * E; c++) {
        // invalid lanes read line 0 (in bounds): lines are transformed independently and
        // their results are not stored, so no branch or select per load
        int n = t + T * c;
        if constexpr (line_direct_in<AX, FWD, KIND>()) {
            bool ng;
            n = pos_elem<N, KIND>(2 * (t + T * (c >> 1)) + (c & 1), ng);
        }
        xin[c] = (line_xt<TIn>)in[ib + (long)n * is];
    }
}

// Transform one line (column) whose inputs are loaded, and store it: L is the line's LDS
// slice
template <typename TIn, typename TOut, typename R, int M, int T, int LINES, int AX, int FWD, int KIND, bool PRE = false>
__device__ __forceinline__ void line_body(const LineArgs &a, const LineAddr &ad, int t, cpx<R> *L,
                                          line_xt<TIn> (&xin)[2 * (M / T)], cpx<R> *lds, long line0)
{
    constexpr int N = 2 * M;
    constexpr int E = M / T;
    constexpr bool SYNC = true;
    TOut *out = reinterpret_cast<TOut *>(a.out);
    const long base = ad.base, stride = ad.stride, tbase = ad.tbase, tstride = ad.tstride;
    const bool valid = ad.valid;
    const int lz = ad.lz;

    // output element n of this thread's line, with the fused output scaling
    auto put = [&](int n, double val) {
        val = a.out_scale * val;
        if constexpr (AX == 0 && !FWD) {
            if (a.plane_add) val += a.plane_add[lz];
            if (a.accumulate) val += (double)out[base + (long)n * stride];
        }
        if (AX == 1 && (a.ztiled & ZT_OUT)) out[tbase + (long)n * tstride] = (TOut)val;
        else out[base + (long)n * stride] = (TOut)val;
    };
    constexpr bool DIRECT_IN = line_direct_in<AX, FWD, KIND>();
    using XT = line_xt<TIn>;
    // staged store (short x lines, line_stage_out): after a barrier, the block writes its lines'
    // outputs -- real-view position store_pos(n) of each line's slice -- as one contiguous run
    constexpr bool SOUT = line_stage_out<TOut, R, T, AX>();
    // x inverse, staged: each line's plane term in a table after the slices (written here,
    // read after the transform's barriers)
    double *planes = reinterpret_cast<double *>(lds + LINES * line_slots<M>());
    if constexpr (SOUT && !FWD) {
        if (a.plane_add && t == 0) planes[threadIdx.x / T] = a.plane_add[lz];
    }
    auto store_staged = [&]() {
        constexpr int LS = line_slots<M>(), NT = T * LINES, C = 2 * E;
        const long rem = a.nlines - line0;
        const int total = (int)(rem < LINES ? rem : LINES) * N;
        const int tid = (int)threadIdx.x;
        TOut *dst = out + line0 * N;
        double val[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int e = tid + NT * c, l2 = e / N, n = e - l2 * N;
            bool neg;
            const int pos = store_pos<N, KIND, FWD>(n, neg);
            val[c] = (e < total) ? (double)lr(lds + l2 * LS, pos) : 0.0;
            if (neg) val[c] = -val[c];
            val[c] = a.out_scale * val[c];
        }
        if constexpr (!FWD) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int e = tid + NT * c;
                if (e < total) {
                    if (a.plane_add) val[c] += planes[e / N];
                    if (a.accumulate) val[c] += (double)dst[e];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int e = tid + NT * c;
            if (e < total) dst[e] = (TOut)val[c];
        }
    };

    // x inverse: the solution being accumulated into and the plane term are read here, with
    // the input, so their latency overlaps the transform instead of sitting before each store
    // (FP32 -> FP64 form only: the FP64 form never accumulates, and holding 2E more doubles
    // there would exceed its register bound)
    constexpr bool XINV = (AX == 0) && !FWD && !std::is_same<TIn, TOut>::value;
    double accv[XINV ? 2 * E : 1];
    double padd = 0.0;
    if constexpr (XINV) {
        if (a.accumulate) {
#pragma unroll
            for (int c = 0; c < 2 * E; c++) accv[c] = (double)out[base + (long)(t + T * c) * stride];
        } else if (a.acc0) {
#pragma unroll
            for (int c = 0; c < 2 * E; c++) {
                const long i = base + (long)(t + T * c) * stride;
                accv[c] = (double)a.acc0[i] + (a.acc1 ? (double)a.acc1[i] : 0.0);
            }
        }
        if (a.plane_add) padd = a.plane_add[lz];
    }

    // ---- twiddle seeds (independent of the data, overlap the load latency) ----
    FFTSeeds<R, M, T> fs;
    fs.load(t, a.tw);
    c64 s1 = {1, 0}, st1 = {1, 0}, s2 = {1, 0}, st2 = {1, 0};
    if constexpr (KIND == KIND_ND || KIND == KIND_DN) {
        s1 = ld64(a.C, t);   // C_k = C_t * D_T^m
        s2 = ld64(a.D, t);   // D_k = D_t * D_T^m
        st1 = ld64(a.D, T);
        st2 = st1;
    } else {
        s1 = ld64(a.A, t);   // A_k = A_t * A_T^m
        st1 = ld64(a.A, T);
        if constexpr (KIND != KIND_PP) {
            s2 = ld64(a.B, t);  // B_k = B_t * B_T^m
            st2 = ld64(a.B, T);
        }
    }

    // PRE: the caller has placed the inputs at their load positions in the line's slice (and
    // passed a barrier), so this load-side processing is skipped
    if constexpr (!PRE) {
        // ---- fused input scaling / Dirichlet boundary term, scatter into LDS ----
        // no table load sits in the element loop except ex[n] on face lines
        const double *ex = a.dtab;
        if constexpr (AX == 0 && FWD) {
            const DirLine dl = dirichlet_line(a, ad.ly, lz);
            const double cyz = dl.cyz, cx0 = dl.cx0, cxn = dl.cxn;
    #pragma unroll
            for (int c = 0; c < 2 * E; c++) {
                xin[c] = (XT)(xin[c] * a.in_scale);  // rhs * c3 * c3 (src/pencilDcmp.cpp:5639)
                xin[c] = (XT)(xin[c] * a.in_scale);
            }
            // one branch per thread, not per element: y/z-face lines, then the x-face elements
            // n = 0 (thread 0, c = 0) and n = N - 1 (thread T - 1, c = 2E - 1)
            if (cyz != 0.0) {
    #pragma unroll
                for (int c = 0; c < 2 * E; c++) xin[c] = (XT)(xin[c] - gld(ex, t + T * c) * cyz);
            }
            if (t == 0) xin[0] = (XT)(xin[0] - cx0);
            if (t == T - 1) xin[2 * E - 1] = (XT)(xin[2 * E - 1] - cxn);
        } else {
    #pragma unroll
            for (int c = 0; c < 2 * E; c++) xin[c] = (XT)(xin[c] * a.in_scale);
        }
        double lsum = 0.0;
        if constexpr (!DIRECT_IN) {
    #pragma unroll
            for (int c = 0; c < 2 * E; c++) {
                const int n = t + T * c;
                const double v = xin[c];
                lsum += v;
                bool neg;
                const int pos = load_pos<N, KIND, FWD>(n, neg);
                lr(L, pos) = neg ? (R)(-v) : (R)v;
            }
        }
        if constexpr (AX == 0 && FWD) {
            if (a.line_sums) {
                // the T threads of a line are adjacent lanes of one wave: butterfly reduction
    #pragma unroll
                for (int off = T / 2; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, T);
                if (t == 0 && valid) a.line_sums[(long)ad.lz * a.ny + ad.ly] = lsum;
            }
        }
        (void)lsum;
        if constexpr (!DIRECT_IN) __syncthreads();
    }

    // ---- pre-processing: FFT input in registers, layout k = t + T*m ----
    cpx<R> v[E];
    if constexpr (DIRECT_IN) {
#pragma unroll
        for (int m = 0; m < E; m++) {
            bool n0, n1;
            pos_elem<N, KIND>(2 * (t + T * m), n0);
            pos_elem<N, KIND>(2 * (t + T * m) + 1, n1);
            v[m] = {n0 ? (R)(-xin[2 * m]) : (R)xin[2 * m], n1 ? (R)(-xin[2 * m + 1]) : (R)xin[2 * m + 1]};
        }
        if constexpr (KIND == KIND_ND || KIND == KIND_DN) {
            c64 cw = s1;
#pragma unroll
            for (int m = 0; m < E; m++) {
                v[m] = c_mul(v[m], to_c<R>(cw));
                cw = m64(cw, st1);
            }
        }
    } else if constexpr (KIND == KIND_ND || KIND == KIND_DN) {
        c64 cw = s1;
#pragma unroll
        for (int m = 0; m < E; m++) {
            v[m] = c_mul(L[padc(t + T * m)], to_c<R>(cw));
            cw = m64(cw, st1);
        }
    } else if constexpr (FWD) {
#pragma unroll
        for (int m = 0; m < E; m++) v[m] = L[padc(t + T * m)];
    } else if constexpr (KIND == KIND_PP) {
        // Z_k = (X_k + conj X_{M-k}) + i e^{2 pi i k/N} (X_k - conj X_{M-k})
        c64 aw = s1;
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            const int k2 = M - k;
            const c64 am = aw;
            aw = m64(aw, st1);
            cpx<R> X1 = {lr(L, k), (k == 0) ? (R)0 : lr(L, N - k)};
            cpx<R> X2 = {lr(L, k2), (k2 == M) ? (R)0 : lr(L, N - k2)};
            cpx<R> s = {X1.x + X2.x, X1.y - X2.y};
            cpx<R> d = {X1.x - X2.x, X1.y + X2.y};
            cpx<R> q = c_mulc(d, to_c<R>(am));
            v[m] = {s.x - q.y, s.y + q.x};
        }
    } else {
        // DCT-III/2 (NN) or DST-III/2 (DD, input reversed):
        // V_k = 1/2 e^{i pi k/(2N)} (X_k - i X_{N-k}),  X_N = 0
        // Z_k = (V_k + conj V_{M-k}) + i e^{2 pi i k/N} (V_k - conj V_{M-k})
        c64 aw = s1, bw = s2;
        const R h = (R)0.70710678118654752440084436210485;
        auto X = [&](int q) -> R {
            if (q >= N) return (R)0;
            if constexpr (KIND == KIND_DD) return lr(L, N - 1 - q);
            else return lr(L, q);
        };
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            const int k2 = M - k;
            const c64 am = aw;
            const cpx<R> b = to_c<R>(bw);
            aw = m64(aw, st1);
            bw = m64(bw, st2);
            // conj(B_{M-k}) = e^{i pi/4} B_k
            const cpx<R> bm = {h * (b.x - b.y), h * (b.x + b.y)};
            cpx<R> V1 = c_scale(c_mulc(cpx<R>{X(k), -X(N - k)}, b), (R)0.5);
            cpx<R> V2 = c_scale(c_mul(cpx<R>{X(k2), -X(N - k2)}, bm), (R)0.5);
            cpx<R> s = {V1.x + V2.x, V1.y - V2.y};
            cpx<R> d = {V1.x - V2.x, V1.y + V2.y};
            cpx<R> q = c_mulc(d, to_c<R>(am));
            v[m] = {s.x - q.y, s.y + q.x};
        }
    }
    if constexpr (!DIRECT_IN) __syncthreads();

    // ---- complex FFT of length M ----
    constexpr bool INV_FFT = (!FWD) && (KIND == KIND_PP || KIND == KIND_NN || KIND == KIND_DD);
    line_fft<R, M, T, INV_FFT ? +1 : -1, SYNC>(v, L, t, fs);

    // ---- post-processing into the line's real LDS view ----
    if constexpr (FWD && (KIND == KIND_PP || KIND == KIND_NN || KIND == KIND_DD)) {
#pragma unroll
        for (int m = 0; m < E; m++) L[padc(t + T * m)] = v[m];
        __syncthreads();
        c64 aw = s1;
        R o1[E], o2[E];
        int p1[E], p2[E];
        if constexpr (KIND == KIND_PP) {
#pragma unroll
            for (int m = 0; m < E; m++) {
                const int k = t + T * m;
                cpx<R> Zk = L[padc(k)];
                cpx<R> Zc = c_conj(L[padc(k == 0 ? 0 : M - k)]);
                // X_k = 1/2 (Zk + Zc) - 1/2 i e^{-2 pi i k/N} (Zk - Zc)
                cpx<R> s = c_scale(c_add(Zk, Zc), (R)0.5);
                cpx<R> q = c_mul(to_c<R>(aw), c_scale(c_sub(Zk, Zc), (R)0.5));
                aw = m64(aw, st1);
                if (k == 0) {
                    o1[m] = Zk.x + Zk.y;
                    p1[m] = 0;
                    o2[m] = Zk.x - Zk.y;
                    p2[m] = M;
                } else {
                    o1[m] = s.x + q.y;
                    p1[m] = k;
                    o2[m] = s.y - q.x;
                    p2[m] = N - k;
                }
            }
        } else {
            c64 bw = s2;
            const R h = (R)0.70710678118654752440084436210485;
#pragma unroll
            for (int m = 0; m < E; m++) {
                const int k = t + T * m;
                cpx<R> Zk = L[padc(k)];
                cpx<R> Zc = c_conj(L[padc(k == 0 ? 0 : M - k)]);
                cpx<R> s = c_scale(c_add(Zk, Zc), (R)0.5);
                cpx<R> q = c_mul(to_c<R>(aw), c_scale(c_sub(Zk, Zc), (R)0.5));
                aw = m64(aw, st1);
                const cpx<R> bk = to_c<R>(bw);
                bw = m64(bw, st2);
                cpx<R> X = {s.x + q.y, s.y - q.x};
                // DCT-II: X_k = 2 Re(e^{-i pi k/(2N)} V_k), X_{N-k} = -2 Im(...)
                if (k == 0) {
                    o1[m] = (R)2 * X.x;  // V_0 real
                    p1[m] = 0;
                    o2[m] = (R)2 * h * (Zk.x - Zk.y);  // 2 Re(e^{-i pi/4} V_M), V_M real
                    p2[m] = M;
                } else {
                    cpx<R> w = c_mul(bk, X);
                    o1[m] = (R)2 * w.x;
                    p1[m] = k;
                    o2[m] = (R)-2 * w.y;
                    p2[m] = N - k;
                }
                if constexpr (KIND == KIND_DD) {
                    p1[m] = N - 1 - p1[m];
                    p2[m] = N - 1 - p2[m];
                }
            }
        }
        if constexpr (SOUT) {
            // short lines: the outputs go to their positions in the slice, then out as one run
            __syncthreads();  // every thread's reads of the spectrum are done
#pragma unroll
            for (int m = 0; m < E; m++) {
                lr(L, p1[m]) = o1[m];
                lr(L, p2[m]) = o2[m];
            }
            __syncthreads();
            store_staged();
            return;
        }
        // outputs k and N - k (N - 1 - k and k - 1 for DD) are contiguous across lanes:
        // stored from registers, no LDS round trip
        if (valid) {
#pragma unroll
            for (int m = 0; m < E; m++) {
                put(p1[m], (double)o1[m]);
                put(p2[m], (double)o2[m]);
            }
        }
        return;
    } else if constexpr (KIND == KIND_ND || KIND == KIND_DN) {
        const R f = FWD ? (R)2 : (R)1;
        c64 dw = s2;
        R o1[E], o2[E];
#pragma unroll
        for (int m = 0; m < E; m++) {
            const int k = t + T * m;
            cpx<R> y = c_mul(v[m], to_c<R>(dw));
            dw = m64(dw, st2);
            o1[m] = f * y.x;
            o2[m] = (KIND == KIND_DN) ? f * y.y : -f * y.y;
            if constexpr (AX == 0) {
                lr(L, 2 * k) = o1[m];
                lr(L, N - 1 - 2 * k) = o2[m];
            }
        }
        if constexpr (AX == 1) {
            // lanes are columns: every output row is a coalesced store from registers
            if (valid) {
#pragma unroll
                for (int m = 0; m < E; m++) {
                    const int k = t + T * m;
                    put(2 * k, (double)o1[m]);
                    put(N - 1 - 2 * k, (double)o2[m]);
                }
            }
            return;
        }
    } else {
        // inverse PP / NN / DD: z_m holds (x_{2m}, x_{2m+1}) or the Makhoul-ordered v
        if constexpr (AX == 1) {
            // lanes are columns: each real-view position p goes to its output row directly
            if (valid) {
#pragma unroll
                for (int m = 0; m < E; m++) {
                    bool n0, n1;
                    const int r0 = pos_elem<N, KIND>(2 * (t + T * m), n0);
                    const int r1 = pos_elem<N, KIND>(2 * (t + T * m) + 1, n1);
                    const double a0 = (double)v[m].x, a1 = (double)v[m].y;
                    put(r0, n0 ? -a0 : a0);
                    put(r1, n1 ? -a1 : a1);
                }
            }
            return;
        }
#pragma unroll
        for (int m = 0; m < E; m++) L[padc(t + T * m)] = v[m];
    }
    __syncthreads();

    if constexpr (SOUT) {
        store_staged();
        return;
    }
    // ---- coalesced store with fused output scaling ----
    if (valid) {
#pragma unroll
        for (int c = 0; c < 2 * E; c++) {
            const int n = t + T * c;
            bool neg;
            const int pos = store_pos<N, KIND, FWD>(n, neg);
            double val = (double)lr(L, pos);
            if (neg) val = -val;
            if constexpr (XINV) {
                val = a.out_scale * val;
                if (a.plane_add) val += padd;
                if (a.accumulate || a.acc0) val += accv[c];
                out[base + (long)n * stride] = (TOut)val;
            } else {
                put(n, val);
            }
        }
    }
}

template <typename TIn, typename TOut, typename R, int M, int T, int LINES, int AX, int FWD, int KIND>
__global__ void __launch_bounds__(T *LINES, (line_min_waves<TIn, TOut, R, M, T, LINES, AX, FWD>())) k_line(LineArgs a0)
{
    LineArgs a = a0;
    int bx, by;
    line_block<TIn, TOut, R, LINES, AX>(a0, a, bx, by);
    constexpr int N = 2 * M;
    constexpr int E = M / T;
    constexpr int LS = line_slots<M>();
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
    const LineAddr ad = line_addr<LINES, AX>(a, N, bx, by, l);
    line_xt<TIn> xin[2 * E];
    const long line0 = (long)bx * LINES;  // x lines: the block's first line
    if constexpr (line_stage_in<TIn, R, T, AX, FWD>()) stage_lines_in<TIn, R, M, T, LINES>(a, line0, lds, t, l, xin);
    else line_load<TIn, R, M, T, AX, FWD, KIND>(a, ad, t, xin);
    line_body<TIn, TOut, R, M, T, LINES, AX, FWD, KIND>(a, ad, t, L, xin, lds, line0);
}

// ---------------------------------------------------------------------------------------
// launch configuration per (R, M, AX)
// ---------------------------------------------------------------------------------------
constexpr int line_threads(int M, int AX)
{
    // E = M / T elements per thread: 8 for power-of-two lengths, 12 with a factor 3, 10 or 20
    // with a factor 5, 14 with a factor 7 (a radix-R pass needs E divisible by R); T = 1
    // below M = 8
    (void)AX;
    switch (M) {
    case 4: return 1;
    case 6: return 1;
    case 8: return 1;
    case 10: return 1;
    case 12: return 1;
    case 14: return 1;
    case 16: return 2;
    case 20: return 2;
    case 24: return 2;
    case 28: return 2;
    case 32: return 4;
    case 40: return 2;
    case 48: return 4;
    case 56: return 4;
    case 64: return 8;
    case 80: return 4;
    case 96: return 8;
    case 112: return 8;
    case 128: return 16;
    case 160: return 8;
    case 192: return 16;
    case 224: return 16;
    case 256: return 32;
    case 320: return 16;
    case 384: return 32;
    case 448: return 32;
    case 512: return 64;
    default: return 0;
    }
}

// The compiled complex lengths M = N / 2 (line length N), radices 2, 3, 5 and 7; every
// dispatch switch over M expands this list
#define PP_FFT_LENGTHS(X) \
    X(4) X(6) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28) X(32) X(40) X(48) X(56) X(64) X(80) X(96) X(112) \
        X(128) X(160) X(192) X(224) X(256) X(320) X(384) X(448) X(512)

// Alternative launch shape of the FP64 transforms (chosen per stage by the autotuner): twice
// the threads per line, half the elements per thread -- fewer registers and more waves in
// flight for one more radix pass (e.g. M = 256: 4 x 4 x 4 x 4 instead of 8 x 8 x 4); offered
// where the plan grows by at most one pass, 0 elsewhere
constexpr int line_threads_alt(int M)
{
    const int T = line_threads(M, 0);
    if (T == 0 || M / T < 8 || !plan_ok(M, M / (2 * T))) return 0;
    return make_plan(M, M / (2 * T)).n <= make_plan(M, M / T).n + 1 ? 2 * T : 0;
}

template <typename R, int M, int AX, int T = line_threads(M, AX)>
constexpr int line_lines()
{
    constexpr int bytes = line_slots<M>() * 2 * (int)sizeof(R);
    // y: 16 FP64 columns (128-byte rows) or up to 32 FP32 columns per block.  FP64 y lines
    // of 1024 points (64 threads each) keep 16 columns in one 1024-thread block (139 KB of
    // LDS, still 16 waves per CU), so their z tiles get whole 128-byte rows; shorter lines
    // with large slices halve the columns instead (more blocks per CU)
    const bool wide = AX == 1 && sizeof(R) == 8 && T >= 64;
    int lines = (AX == 0) ? 256 / T : (sizeof(R) == 8 ? 16 : 32);
    if (lines < 1) lines = 1;
    if (lines > 64) lines = 64;
    const int budget = (AX == 0) ? 40 * 1024 : (wide ? 144 * 1024 : 80 * 1024);
    while (lines > 1 && lines * bytes > budget) lines /= 2;
    while (lines > 1 && lines * T > (wide ? 1024 : 512)) lines /= 2;
    return lines;
}

// FP32 y transforms of length 2M have a 32-column block (1024 threads at most, 144 KB of LDS)
template <int M, int T = line_threads(M, 1)>
constexpr bool line_tiled32()
{
    return T > 0 && 32 * T <= 1024 && 32 * line_slots<M>() * 2 * (int)sizeof(float) <= 144 * 1024;
}

// ... only in the shape with half the threads per column (and twice the elements)
template <int M, int T = line_threads(M, 1)>
constexpr bool tiled32_half()
{
    return T > 0 && !line_tiled32<M, T>() && T % 2 == 0 && line_tiled32<M, T / 2>() && plan_ok(M, 2 * (M / T));
}

template <typename TIn, typename TOut, typename R, int M, int T, int LINES, int AX, int FWD, int KIND>
hipError_t launch_line_blocks(const LineArgs &a, hipStream_t s)
{
    const size_t lds = (size_t)LINES * line_slots<M>() * sizeof(cpx<R>) +
                       ((line_stage_out<TOut, R, T, AX>() && !FWD) ? LINES * sizeof(double) : 0);
    auto kern = k_line<TIn, TOut, R, M, T, LINES, AX, FWD, KIND>;
    if (lds > 64 * 1024) {
        static bool attr_set = false;
        if (!attr_set) {
            hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            attr_set = true;
        }
    }
    dim3 block(T * LINES);
    dim3 grid;
    if (a.items) grid = dim3((unsigned)a.nlines);  // batched: nlines holds the item count
    else if (AX == 0) grid = dim3((unsigned)((a.nlines + LINES - 1) / LINES));
    else grid = dim3((unsigned)((a.nlines + LINES - 1) / LINES), (unsigned)a.nplanes);
    hipLaunchKernelGGL(kern, grid, block, lds, s, a);
    return hipGetLastError();
}

template <typename TIn, typename TOut, typename R, int M, int AX, int FWD, int KIND, int V = 0>
hipError_t launch_line_cfg(const LineArgs &a, hipStream_t s)
{
    constexpr int T = V ? line_threads_alt(M) : line_threads(M, AX);
    if constexpr (T == 0) {
        return hipErrorInvalidValue;
    } else {
        // y keeps the default's columns per block in either shape (the z-tile width)
        constexpr int LINES = (V && AX == 1) ? line_lines<R, M, AX>() : line_lines<R, M, AX, T>();
        // the 16-column FP64 blocks of 1024-point y lines (one block per CU) pay off where they
        // write or read whole z-tile rows (two 8-column blocks per CU, each writing half of
        // every tile row, and the same block loading a second plane's columns while it
        // transforms the first both measured slower at 1024^3); on the field in place two
        // 8-column blocks per CU run faster
        if constexpr (AX == 1 && sizeof(R) == 8 && T >= 64 && LINES == 16) {
            if (!a.ztiled && !a.items) return launch_line_blocks<TIn, TOut, R, M, T, LINES / 2, AX, FWD, KIND>(a, s);
        }
        // FP32 z tiles are 32 columns wide (128-byte rows): 32-column blocks of up to 1024
        // threads write or read whole tile rows
        if constexpr (AX == 1 && sizeof(R) == 4 && line_tiled32<M, T>()) {
            if (a.ztiled && a.tile_w == 32 && !a.items) {
                // alternative: half the threads per column, twice the elements (one radix pass
                // and one LDS exchange fewer, 512-thread blocks)
                if constexpr (T % 2 == 0 && plan_ok(M, 2 * (M / T))) {
                    if (a.variant) return launch_line_blocks<TIn, TOut, R, M, T / 2, 32, AX, FWD, KIND>(a, s);
                }
                return launch_line_blocks<TIn, TOut, R, M, T, 32, AX, FWD, KIND>(a, s);
            }
        } else if constexpr (AX == 1 && sizeof(R) == 4 && tiled32_half<M, T>()) {
            // lines too long for 32 columns of the default shape (1024 points): the
            // half-thread shape is the 32-column block
            if (a.ztiled && a.tile_w == 32 && !a.items)
                return launch_line_blocks<TIn, TOut, R, M, T / 2, 32, AX, FWD, KIND>(a, s);
        }
        return launch_line_blocks<TIn, TOut, R, M, T, LINES, AX, FWD, KIND>(a, s);
    }
}

template <typename TIn, typename TOut, typename R, int AX, int FWD, int KIND, int V = 0>
hipError_t launch_line_sz(int N, const LineArgs &a, hipStream_t s)
{
    switch (N / 2) {
#define PPM(m) \
    case m: return launch_line_cfg<TIn, TOut, R, m, AX, FWD, KIND, V>(a, s);
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return hipErrorInvalidValue;
    }
}

// Precision forms of one transform kind (instantiated once per kind in pp_line_<kind>.hip):
//   F64        : FP64 in/out/compute, every axis and direction
//   F32        : FP32 in/out/compute, every axis and direction
//   F64_TO_F32 : x forward reading the FP64 right-hand side, FP32 compute and output
//   F32_TO_F64 : x inverse writing (or accumulating into) the FP64 solution
enum LineForm : int { FORM_F64 = 0, FORM_F32 = 1, FORM_F64_TO_F32 = 2, FORM_F32_TO_F64 = 3 };

template <int KIND>
hipError_t launch_line_kind(int form, int N, int AX, int FWD, const LineArgs &a, hipStream_t s)
{
    switch (form) {
    case FORM_F64:
        if (a.variant) {
            if (AX == 0) return FWD ? launch_line_sz<double, double, double, 0, 1, KIND, 1>(N, a, s)
                                    : launch_line_sz<double, double, double, 0, 0, KIND, 1>(N, a, s);
            return FWD ? launch_line_sz<double, double, double, 1, 1, KIND, 1>(N, a, s)
                       : launch_line_sz<double, double, double, 1, 0, KIND, 1>(N, a, s);
        }
        if (AX == 0) return FWD ? launch_line_sz<double, double, double, 0, 1, KIND>(N, a, s)
                                : launch_line_sz<double, double, double, 0, 0, KIND>(N, a, s);
        return FWD ? launch_line_sz<double, double, double, 1, 1, KIND>(N, a, s)
                   : launch_line_sz<double, double, double, 1, 0, KIND>(N, a, s);
    case FORM_F32:
        if (AX == 0) return FWD ? launch_line_sz<float, float, float, 0, 1, KIND>(N, a, s)
                                : launch_line_sz<float, float, float, 0, 0, KIND>(N, a, s);
        return FWD ? launch_line_sz<float, float, float, 1, 1, KIND>(N, a, s)
                   : launch_line_sz<float, float, float, 1, 0, KIND>(N, a, s);
    case FORM_F64_TO_F32:
        if (AX == 0 && FWD) return launch_line_sz<double, float, float, 0, 1, KIND>(N, a, s);
        return hipErrorInvalidValue;
    case FORM_F32_TO_F64:
        if (AX == 0 && !FWD) return launch_line_sz<float, double, float, 0, 0, KIND>(N, a, s);
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

// a compiled Stockham plan serves length n (other lengths: pp_line_bs.h)
inline bool line_compiled(int n) { return n % 2 == 0 && line_threads(n / 2, 0) > 0; }

// the FP64 transforms of length N have an alternative launch shape
inline bool line_has_variant(int n) { return line_compiled(n) && line_threads_alt(n / 2) > 0; }

// the FP32 y transform of length N can write / read 32-column z tiles in whole rows
inline bool line_tiled32_f32(int N)
{
    switch (line_compiled(N) ? N / 2 : 0) {
#define PPM(m) \
    case m: return line_tiled32<m>() || tiled32_half<m>();
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return false;
    }
}

template <int M>
constexpr bool tiled32_has_variant()
{
    constexpr int T = line_threads(M, 1);
    return line_tiled32<M>() && T % 2 == 0 && plan_ok(M, 2 * (M / (T > 0 ? T : 1)));
}

// ... and has the alternative shape (half the threads per column) for those tiles
inline bool line_tiled32_f32_variant(int N)
{
    switch (line_compiled(N) ? N / 2 : 0) {
#define PPM(m) \
    case m: return tiled32_has_variant<m>();
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return false;
    }
}

// lines (x) or columns (y) per block of the launch for line length N, precision R
template <typename R>
inline int line_block_lines(int N, int AX)
{
    switch (N / 2) {
#define PPM(m) \
    case m: return AX == 0 ? line_lines<R, m, 0>() : line_lines<R, m, 1>();
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return 0;
    }
}

template <int M>
constexpr int alt_lines_x()
{
    if constexpr (line_threads_alt(M) > 0) return line_lines<double, M, 0, line_threads_alt(M)>();
    else return 0;
}

// lines (x) or columns (y) per block of the alternative FP64 shape (y keeps the default's)
inline int line_block_lines_alt(int N, int AX)
{
    if (AX == 1) return line_block_lines<double>(N, 1);
    switch (N / 2) {
#define PPM(m) \
    case m: return alt_lines_x<m>();
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return 0;
    }
}

}  // namespace pp
