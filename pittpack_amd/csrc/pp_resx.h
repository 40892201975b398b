// pp_resx.h -- the mixed form's correction step in one kernel: the FP64 residual of the
// dz^2-scaled 7-point operator, r = (rhs*dz^2 - Dirichlet term) - L u, computed inside the FP32
// x-forward transform of the lines it feeds.  r never goes through memory: against the
// residual kernel followed by the transform (pp_mixed.hip k_residual_pair, then k_line) this
// moves f, u and the transformed lines only, 8 bytes per grid point fewer.
//
// Replaces, for the refinement's corrections, the reference's right-hand-side assembly ahead
// of the x transform (assignRhs's dz^2 scaling, src/pencilDcmp.cpp:5591-5647, and
// modifyRhsDirichlet, 4623-4835) with the residual of the same system.
//
// A block owns LINES consecutive x lines = rows y0 .. y0+LINES-1 of one plane z (the host
// launches it only when ny is a multiple of LINES).  It stages u(z) of those rows and of the
// two rows around them (ghost rows at the y faces) in LDS in FP64, reads u(z-1), u(z+1) and f
// for its own cells, computes r four cells per lane with the residual kernel's arithmetic,
// rounds it to FP32 into the lines' slices and hands them to the transform (line_body).
#pragma once
#include <hip/hip_runtime.h>

#include "pp_line.h"
#include "pp_mixed.h"

namespace pp {

typedef double rx_d2 __attribute__((ext_vector_type(2)));
typedef double rx_d4 __attribute__((ext_vector_type(4)));
typedef float rx_f4 __attribute__((ext_vector_type(4)));

// bytes of LDS of the fused kernel: the FP64 row tile, or the lines' slices (they alias)
template <int M, int LINES>
constexpr size_t resx_lds()
{
    // the slices overwrite the tile once it is read (40 KB for 512-point lines: four blocks
    // per CU); the plane-0 wave sums go to the tile's tail, which the slices never reach
    const size_t tile = (size_t)(LINES + 2) * 2 * M * sizeof(double);
    const size_t slices = (size_t)LINES * line_slots<M>() * sizeof(cpx<float>);
    return tile > slices ? tile : slices;
}

// Register bound: the waves per SIMD the LDS tile allows (blocks per CU x waves per block / 4):
// 4 for 512-point lines (40 KB per 256-thread block, <= 128 VGPRs), 3 for 1024-point ones (48
// KB); at most 3 for shorter lines and for the batched form (their second-step kernels spill at
// 128) and at most 2 where a thread holds other than 8 elements (their transforms need more
// registers)
template <int M, int T, int LINES, bool BATCH>
constexpr int resx_waves()
{
    const int blocks = (int)((160 * 1024) / resx_lds<M, LINES>());
    int w = blocks * (T * LINES / 64) / 4;
    if ((M / T) != 8 && w > 2) w = 2;
    if ((M < 256 || BATCH) && w > 3) w = 3;
    return w < 1 ? 1 : (w > 4 ? 4 : w);
}

template <int M, int T, int LINES, int KIND, int SRC, bool BATCH>
__global__ void __launch_bounds__(T *LINES, (resx_waves<M, T, LINES, BATCH>())) k_resx(LineArgs a0, ResArgs r0,
                                                                                       const ResArgs *rsolves)
{
    using R = float;
    constexpr int N = 2 * M, E = M / T, LS = line_slots<M>(), NT = T * LINES;
    // four adjacent x cells per lane and access: u parts as one 16-byte load each, f as two
    constexpr int QR = N / 4;            // cell quads per row
    constexpr int QQ = LINES * QR / NT;  // own quads per thread
    static_assert(N % 4 == 0 && QQ * NT == LINES * QR && 2 * QQ == E, "a block's rows are E / 2 quads per thread");
    static_assert(NT % 64 == 0, "whole waves");
    constexpr int HC = (2 * QR + NT - 1) / NT;  // quads of the two outer rows per thread
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cpx<R> *lds = reinterpret_cast<cpx<R> *>(smem_raw);
    double *ut = reinterpret_cast<double *>(smem_raw);  // (LINES + 2) rows of N: rows y0-1 .. y0+LINES
    static_assert((size_t)LINES * line_slots<M>() + NT / 64 <= (size_t)(LINES + 2) * N,  // in doubles
                  "the wave sums sit in the tile's tail, past the slices");
    double *red = ut + (LINES + 2) * N - NT / 64;  // plane-0 wave sums, once the tile is read
    const int tid = (int)threadIdx.x, t = tid % T, l = tid / T;
    LineArgs a = a0;
    ResArgs r = r0;
    int g, z;
    if constexpr (BATCH) {
        // batched: block b works on items[b] = {member, x-line block, 0}; the member's line
        // fields resolve into a (line_block) and its residual arguments come from rsolves
        int ib, jb;
        line_block<float, float, float, LINES, 0>(a0, a, ib, jb);
        r = rsolves[a0.items[blockIdx.x].solve];
        const int ngm = r.ny / LINES;
        z = ib / ngm;
        g = ib % ngm;
    } else {
        // Block numbering: with ng a multiple of 8, XCD k (blocks b = k mod 8) owns row groups
        // k*ng/8 .. (k+1)*ng/8 - 1 of every plane and takes them plane by plane, so the z- and
        // y-neighbour rows a block reads were read by blocks that ran on the same XCD just
        // before (its own L2); otherwise groups in plane order
        const int ngs = r.ny / LINES;
        if (ngs % 8 == 0) {
            const int gr = ngs / 8, k = (int)blockIdx.x % 8, s = (int)blockIdx.x / 8;
            z = s / gr;
            g = k * gr + s % gr;
        } else {
            z = (int)blockIdx.x / ngs;
            g = (int)blockIdx.x % ngs;
        }
    }
    const int ny = r.ny, nz = r.nz, ng = ny / LINES;
    const int y0 = g * LINES, bx = z * ng + g;
    const long sy = N, sz = (long)N * ny;
    const long p0 = (long)z * sz + (long)y0 * sy;  // first cell of the block's rows
    using gd2 = const __attribute__((address_space(1))) rx_d2;
    using gf4 = const __attribute__((address_space(1))) rx_f4;
    // (double)u0 + (double)u1 per cell, as the residual kernel sums the parts
    auto sum4 = [](rx_f4 a0, rx_f4 a1) -> rx_d4 {
        rx_d4 v = __builtin_convertvector(a0, rx_d4);
        if constexpr (SRC == RES_U0_U1) v += __builtin_convertvector(a1, rx_d4);
        return v;
    };
    auto ldu4 = [&](long o) -> rx_d4 {
        const rx_f4 p = *(gf4 *)(r.u0 + o);
        rx_f4 q = p;
        if constexpr (SRC == RES_U0_U1) q = *(gf4 *)(r.u1 + o);
        return sum4(p, q);
    };

    // ---- loads: own rows of u(z), the outer rows, u(z -/+ 1) and f of the own cells ----
    // outer rows: inside the grid the neighbouring rows; at a y face the ghost row (P: the
    // wrapped row, N: the face row itself, D: its negative)
    int rowm = y0 - 1, rowp = y0 + LINES;
    double sgm = 1.0, sgp = 1.0;
    if (y0 == 0) {
        rowm = r.bc[2] == 'P' ? ny - 1 : 0;
        sgm = r.bc[2] == 'D' ? -1.0 : 1.0;
    }
    if (y0 + LINES == ny) {
        rowp = r.bc[3] == 'P' ? 0 : ny - 1;
        sgp = r.bc[3] == 'D' ? -1.0 : 1.0;
    }
    // z neighbours: inside the grid the next planes; at a z face P wraps, N and D read this
    // plane's cell (signed): every quad loads unconditionally, so no branch holds a wait
    const bool zlo = z == 0, zhi = z == nz - 1;
    const long om = zlo ? (r.bc[4] == 'P' ? (long)(nz - 1) * sz : 0) : -sz;
    const long op = zhi ? (r.bc[5] == 'P' ? -(long)(nz - 1) * sz : 0) : sz;
    const double sgzm = (zlo && r.bc[4] == 'D') ? -1.0 : 1.0, sgzp = (zhi && r.bc[5] == 'D') ? -1.0 : 1.0;
    {
        // own rows (kept in the tile only) and outer rows: loaded, then written to the tile
        rx_d4 uc[QQ], hv[HC];
#pragma unroll
        for (int c = 0; c < QQ; c++) uc[c] = ldu4(p0 + 4 * (tid + NT * c));
#pragma unroll
        for (int i = 0; i < HC; i++) {
            const int h = min(tid + NT * i, 2 * QR - 1), j = h < QR ? 0 : 1;
            hv[i] = ldu4((long)z * sz + (long)(j ? rowp : rowm) * sy + 4 * (h - j * QR));
        }
#pragma unroll
        for (int c = 0; c < QQ; c++) {
            double *d = ut + N + 4 * (tid + NT * c);
            *(rx_d2 *)d = uc[c].lo;
            *(rx_d2 *)(d + 2) = uc[c].hi;
        }
#pragma unroll
        for (int i = 0; i < HC; i++) {
            const int h = tid + NT * i, j = h < QR ? 0 : 1;
            if (h < 2 * QR) {
                double *d = ut + (j ? (LINES + 1) * N : 0) + 4 * (h - j * QR);
                const rx_d4 v = hv[i] * (j ? sgp : sgm);
                *(rx_d2 *)d = v.lo;
                *(rx_d2 *)(d + 2) = v.hi;
            }
        }
    }
    // z neighbours and the right-hand side of the own quads, as raw loads issued PD quads ahead
    // of their use: each quad's arithmetic sits between scheduling barriers, so only the quads
    // in flight hold registers
    constexpr int PD = 1;
    struct Raw {
        rx_f4 m0, m1, p0, p1;
        rx_d2 f0, f1, x0, x1;  // f and, with Dirichlet faces, the ex profile of the quad
        double ey;             // and ey of its row
    };
    const int dm = r.dmask;
    Raw raw[QQ];
    auto issue = [&](int c) {
        const long o = p0 + 4 * (tid + NT * c);
        raw[c].m0 = *(gf4 *)(r.u0 + o + om);
        raw[c].p0 = *(gf4 *)(r.u0 + o + op);
        if constexpr (SRC == RES_U0_U1) {
            raw[c].m1 = *(gf4 *)(r.u1 + o + om);
            raw[c].p1 = *(gf4 *)(r.u1 + o + op);
        } else {
            raw[c].m1 = raw[c].m0;
            raw[c].p1 = raw[c].p0;
        }
        raw[c].f0 = *(gd2 *)(r.f + o);
        raw[c].f1 = *(gd2 *)(r.f + o + 2);
        // Dirichlet profiles (cached table loads; every quad loads them when a face is D, so
        // the loop has no branch)
        if (dm) {
            const int q = 4 * (tid + NT * c), l2 = q / N, x = q - l2 * N;
            raw[c].x0 = *(gd2 *)(r.dtab + x);
            raw[c].x1 = *(gd2 *)(r.dtab + x + 2);
            raw[c].ey = *(const __attribute__((address_space(1))) double *)(r.dtab + N + y0 + l2);
        } else {
            raw[c].x0 = raw[c].x1 = rx_d2{0.0, 0.0};
            raw[c].ey = 0.0;
        }
    };
#pragma unroll
    for (int c = 0; c < PD && c < QQ; c++) issue(c);
    __syncthreads();

    // ---- residual per cell quad (the residual kernel's arithmetic) ----
    const double *ef = r.dtab + N + ny + nz;
    const double ezz = r.dtab[N + ny + z];
    const double ef0 = ef[0], ef1 = ef[1], ef2 = ef[2], ef3 = ef[3], ef4 = ef[4], ef5 = ef[5];
    const double defect = r.pin_defect ? r.pin_defect[0] : 0.0;
    const char bx0 = r.bc[0], bx1 = r.bc[1];
    const double gxm = bx0 == 'D' ? -1.0 : 1.0, gxp = bx1 == 'D' ? -1.0 : 1.0;
    // pinned (0,0) column: the block's partial sum of u over plane 0 (periodic z: planes 0 and
    // nz - 1, the latter held as u(z-1)); ng entries, one per row group
    const bool sums = r.upart && zlo, zwrap = r.bc[4] == 'P';
    double sp = 0.0;
    rx_f4 rv[QQ];
#pragma unroll
    for (int c = 0; c < QQ; c++) {
        if (c + PD < QQ) issue(c + PD);
        __builtin_amdgcn_sched_barrier(0);
        const int q = 4 * (tid + NT * c), l2 = q / N, x = q - l2 * N, y = y0 + l2;
        const double *row = ut + (l2 + 1) * N;
        const rx_d2 ua = *(const rx_d2 *)(row + x), ub = *(const rx_d2 *)(row + x + 2);
        const rx_d2 ma = *(const rx_d2 *)(row - N + x), mb = *(const rx_d2 *)(row - N + x + 2);
        const rx_d2 pa = *(const rx_d2 *)(row + N + x), pb = *(const rx_d2 *)(row + N + x + 2);
        const double u[4] = {ua.x, ua.y, ub.x, ub.y};
        const double uym[4] = {ma.x, ma.y, mb.x, mb.y}, uyp[4] = {pa.x, pa.y, pb.x, pb.y};
        const rx_d4 umv = sum4(raw[c].m0, raw[c].m1) * sgzm, upv = sum4(raw[c].p0, raw[c].p1) * sgzp;
        const double um[4] = {umv.x, umv.y, umv.z, umv.w}, up[4] = {upv.x, upv.y, upv.z, upv.w};
        const double fv[4] = {raw[c].f0.x, raw[c].f0.y, raw[c].f1.x, raw[c].f1.y};
        if (sums) sp += ((u[0] + u[1]) + (u[2] + u[3])) + (zwrap ? (um[0] + um[1]) + (um[2] + um[3]) : 0.0);
        // x neighbours: inside the row from the tile; at x = 0 and x = N - 1 the ghost closure
        // (P wraps within the row, N mirrors: the cell itself, D: its negative)
        const int il = x > 0 ? x - 1 : (bx0 == 'P' ? N - 1 : 0);
        const int ir = x + 4 < N ? x + 4 : (bx1 == 'P' ? 0 : N - 1);
        const double uxl = (x > 0 ? 1.0 : gxm) * row[il];
        const double uxr = (x + 4 < N ? 1.0 : gxp) * row[ir];
        // Dirichlet profiles of the quad's cells and row
        const double exv[4] = {raw[c].x0.x, raw[c].x0.y, raw[c].x1.x, raw[c].x1.y};
        const double eyc = raw[c].ey;
        const bool fy0 = (dm & 4) && y == 0, fy1 = (dm & 8) && y == ny - 1;
        float rq[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double xl = i == 0 ? uxl : u[i - 1], xr = i == 3 ? uxr : u[i + 1];
            const double Lu = ((xr - u[i]) - (u[i] - xl)) * r.Cx + ((uyp[i] - u[i]) - (u[i] - uym[i])) * r.Cy +
                              ((up[i] - u[i]) - (u[i] - um[i]));
            // Dirichlet terms, as the residual kernel forms them, with selects (no branch per
            // cell): x and y faces a per-cell constant times ez[z], z faces a constant on the
            // end planes
            double cxy = 0.0;
            if (i == 0) cxy = ((dm & 1) && x == 0) ? cxy + 2. * ef0 * eyc : cxy;
            if (i == 3) cxy = ((dm & 2) && x + 3 == N - 1) ? cxy + 2. * ef1 * eyc : cxy;
            cxy = fy0 ? cxy + 2. * exv[i] * ef2 : cxy;
            cxy = fy1 ? cxy + 2. * exv[i] * ef3 : cxy;
            const double czlo = (dm & 16) ? 2. * exv[i] * eyc * ef4 : 0.0;
            const double czhi = (dm & 32) ? 2 * exv[i] * eyc * ef5 : 0.0;
            double v = fv[i];
            v = v * r.c3;
            v = v * r.c3;
            v = v - cxy * ezz;
            if (zlo) v = v - czlo;
            if (zhi) v = v - czhi;
            double res = v - Lu;
            if (zlo) res += defect;
            rq[i] = (float)res;
        }
        // the result is pinned here (an empty asm that takes it in registers): otherwise the
        // arithmetic sinks to the slice writes after the barrier and every quad's loads stay live
        asm volatile("" : "+v"(rq[0]), "+v"(rq[1]), "+v"(rq[2]), "+v"(rq[3]));
        rv[c] = rx_f4{rq[0], rq[1], rq[2], rq[3]};
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();  // the tile is read: the slices and the wave sums overwrite it
    if (sums) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sp += __shfl_down(sp, off);
        if ((tid & 63) == 0) red[tid >> 6] = sp;
        __syncthreads();
    }
    if (sums && tid == 0) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NT / 64; i++) s += red[i];
        r.upart[g] = s;
    }

    const LineAddr ad = line_addr<LINES, 0>(a, N, bx, 0, l);
    cpx<R> *L = lds + l * LS;
    line_xt<float> xin[2 * E];
    // no line sums wanted: r goes straight to the transform's input positions in the slices
    // (the scatter line_body would do), and the transform starts from there (first step 0.52
    // -> 0.48 ms at 512^3); single-solve kernels of 8 elements per thread only, where the
    // second transform path still fits the register bound
    constexpr bool DIRECT = (M / T == 8) && !BATCH;
    if (DIRECT && !a.line_sums) {
#pragma unroll
        for (int c = 0; c < QQ; c++) {
            const int q = 4 * (tid + NT * c), l2 = q / N, x = q - l2 * N;
            const float rq[4] = {rv[c].x, rv[c].y, rv[c].z, rv[c].w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                bool neg;
                const int pos = load_pos<N, KIND, 1>(x + i, neg);
                lr(lds + l2 * LS, pos) = neg ? -rq[i] : rq[i];
            }
        }
        __syncthreads();
        if constexpr (DIRECT) line_body<float, float, float, M, T, LINES, 0, 1, KIND, true>(a, ad, t, L, xin, lds, (long)bx * LINES);
        return;
    }
    // ---- r into the lines' slices, then each thread's inputs n = t + T c of its line (whose
    // line sums the transform forms) ----
#pragma unroll
    for (int c = 0; c < QQ; c++) {
        const int q = 4 * (tid + NT * c), l2 = q / N, x = q - l2 * N;
        lr(lds + l2 * LS, x) = rv[c].x;
        lr(lds + l2 * LS, x + 1) = rv[c].y;
        lr(lds + l2 * LS, x + 2) = rv[c].z;
        lr(lds + l2 * LS, x + 3) = rv[c].w;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2 * E; c++) xin[c] = lr(L, t + T * c);
    __syncthreads();  // the slices are rewritten by the pre-processing next
    line_body<float, float, float, M, T, LINES, 0, 1, KIND>(a, ad, t, L, xin, lds, (long)bx * LINES);
}

template <int M, int KIND, int SRC>
hipError_t launch_resx_cfg(const LineArgs &a, const ResArgs &r, const ResArgs *rsolves, hipStream_t s)
{
    constexpr int T = line_threads(M, 0);
    if constexpr (T == 0) {
        return hipErrorInvalidValue;
    } else {
        constexpr int LINES = line_lines<float, M, 0>();
        // batched (a.items, a.nlines items): every member's nx = 2M and ny % LINES == 0 is the
        // planner's to ensure; one solve: checked here
        if (!a.items && (r.nx != 2 * M || r.ny % LINES != 0 || a.nlines != (long)r.ny * r.nz)) return hipErrorInvalidValue;
        if (a.items && !rsolves) return hipErrorInvalidValue;
        constexpr size_t lds = resx_lds<M, LINES>();
        // one solve and the batched form (members resolved per block) are separate kernels, so
        // the single solve's kernel carries no per-member indirection
        const int bat = a.items ? 1 : 0;
        auto kern = bat ? k_resx<M, T, LINES, KIND, SRC, true> : k_resx<M, T, LINES, KIND, SRC, false>;
        if constexpr (lds > 64 * 1024) {
            static bool attr_set[2] = {false, false};
            if (!attr_set[bat]) {
                hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e != hipSuccess) return e;
                attr_set[bat] = true;
            }
        }
        const unsigned blocks = bat ? (unsigned)a.nlines : (unsigned)((long)(r.ny / LINES) * r.nz);
        if (blocks == 0) return hipSuccess;
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(T * LINES), lds, s, a, r, rsolves);
        return hipGetLastError();
    }
}

// the fused correction step of one x transform kind (instantiated in pp_line_<kind>.hip):
// a is the FP32 x-forward's arguments (out: the work field, line sums), r the residual's
// (src: RES_U0 or RES_U0_U1); batched (a.items): rsolves holds each member's residual
// arguments (device memory) and r is unused
template <int KIND>
hipError_t launch_resx_kind(int N, int src, const LineArgs &a, const ResArgs &r, const ResArgs *rsolves,
                            hipStream_t s)
{
    if (src != RES_U0 && src != RES_U0_U1) return hipErrorInvalidValue;
    switch (N / 2) {
#define PPM(m) \
    case m: return src == RES_U0 ? launch_resx_cfg<m, KIND, RES_U0>(a, r, rsolves, s) \
                                 : launch_resx_cfg<m, KIND, RES_U0_U1>(a, r, rsolves, s);
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return hipErrorInvalidValue;
    }
}

// x lines per block of the fused kernel for line length N (0: no compiled plan); the fused
// step applies where ny is a multiple of it, and writes ny / resx_lines(N) partial sums
inline int resx_lines(int N)
{
    if (N % 2) return 0;
    switch (N / 2) {
#define PPM(m) \
    case m: return line_threads(m, 0) ? line_lines<float, m, 0>() : 0;
        PP_FFT_LENGTHS(PPM)
#undef PPM
    default: return 0;
    }
}

// dispatch over the x transform kind (pp_solver.hip)
hipError_t launch_resx(int kind, int N, int src, const LineArgs &a, const ResArgs &r, const ResArgs *rsolves,
                       hipStream_t s);

}  // namespace pp
