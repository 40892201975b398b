// pp_launch.h -- host launchers shared by the single-solve handle and the batched form.
#pragma once
#include <hip/hip_runtime.h>

#include "pp_line.h"
#include "pp_zsolve.h"

namespace pp {

// line transform of one kind (pp_line_<kind>.hip) in precision form `form` (LineForm)
hipError_t launch_line(int form, int kind, int N, int AX, int FWD, const LineArgs &a, hipStream_t s);

// z solves (pp_zsolve_pcr.hip, pp_zsolve_scan.hip).  FP64 columns hold at most 32 rows per
// lane (nz <= 512), FP32 columns at most 64 (nz <= 1024).
hipError_t launch_zsolve_pcr_f64(int nz, const ZArgs &a, hipStream_t s);
hipError_t launch_zsolve_scan_f64(int nz, const ZArgs &a, hipStream_t s);
bool zsolve_supported_f64(int nz);
// FP64 columns of 513..1024 rows (pp_zsolve_long.hip): in place, not tiled, every column
// written back; launch_zsolve_pcr_f64 forwards these lengths here
bool zsolve_long_supported_f64(int nz);
hipError_t launch_zsolve_long_f64(int nz, const ZArgs &a, hipStream_t s);
// multigrid alternative (pp_zmg.hip): V-cycles until max|residual| <= tol * max|rhs|
hipError_t launch_zsolve_mg_f64(const ZArgs &a, int max_cycles, double tol, hipStream_t s);
// the reference's alternatives (pp_zalt.hip): full parallel cyclic reduction and cyclic
// reduction, one wave per column, nz <= 1024
bool zsolve_alt_supported(int nz);
hipError_t launch_zsolve_pcr_full_f64(const ZArgs &a, hipStream_t s);
hipError_t launch_zsolve_cr_f64(const ZArgs &a, hipStream_t s);
hipError_t launch_zsolve_pcr_f32(int nz, const ZArgs &a, hipStream_t s);
bool zsolve_supported_f32(int nz);
// FP64 re-solve of the near-singular columns of FP32 data, one block per listed 16-column
// tile (a.items over a.solves); columns up to 512 rows
bool zfix_tiles_supported(int nz);
hipError_t launch_zfix_tiles_f32(int nz_max, const ZArgs &a, hipStream_t s);

// the reference's Thomas sweeps (thomasLowMem) on listed FP64 columns in place or on z tiles
// (a's layout), one wave per column, nz <= 1024.  solve >= 0: batched, the member's field and
// sizes from a.solves; toff: offset of the column's factors {1 / bet_j, tmp_j} in tab
// (zthomas_factors, 2 nz values)
struct ZThomasItem {
    int sx, sy, solve, pad;
    long toff;
};
hipError_t launch_zthomas_f64(const ZArgs &a, const ZThomasItem *items, const double *tab, int nitems, hipStream_t s);

// thomasLowMem's factors of one column (src/triDiag.cpp:1693-1822): out[j] = 1 / bet_j,
// out[n + j] = tmp_j (tmp_0 unused), with b0 / bn the end rows' diagonals
inline void zthomas_factors(double eig, double b0, double bn, int n, double *out)
{
    double bet = b0;
    out[0] = 1.0 / bet;
    out[n] = 0.0;
    for (int j = 1; j < n; j++) {
        const double tmp = 1.0 / bet;
        bet = (j == n - 1 ? bn : eig) - tmp;
        out[j] = 1.0 / bet;
        out[n + j] = tmp;
    }
}

// periodic z (pp_zperiodic.hip): circulant columns by z FFT, FP64 arithmetic on FP64 or FP32
// storage; the singular column by the reference's own branch
struct ZPerArgs {
    void *data;  // [nz][SY][SX], in place; double (f64 launchers) or float (f32 launchers)
    int SX, SY, nz;
    const double *lamx, *lamy;
    double Cx, Cy;
    const double *tw;
    const double *c2;
    // mixed form: the singular column is written as zeros (its contribution is solved from
    // plane sums in FP64 elsewhere); otherwise it is left for the reference's branch
    int pinned_zero;
    // batched launch of the one-pass kernel (null: one solve, grid = column blocks x SY):
    // block b works on columns bx * zper_lines(nz) ... of row sy of solve items[b].solve
    // (items[b] = {solve, bx, sy}); every listed solve has length nz
    const ZItem *items;
    const ZSolve *solves;
};
// one-pass kernel: nz with a compiled complex plan (zperiodic_supported); other nz: real
// z-line transforms around launch_zper_divide
bool zperiodic_supported(int nz);
int zper_lines(int nz);  // columns per block of the one-pass kernel
hipError_t launch_zperiodic_f64(const ZPerArgs &a, hipStream_t s);
hipError_t launch_zperiodic_f32(const ZPerArgs &a, hipStream_t s);
// batched launches: grid = nitems
hipError_t launch_zperiodic_f64_batch(const ZPerArgs &a, int nitems, hipStream_t s);
hipError_t launch_zperiodic_f32_batch(const ZPerArgs &a, int nitems, hipStream_t s);
hipError_t launch_zper_divide(const ZPerArgs &a, hipStream_t s);
hipError_t launch_zper_divide_f32(const ZPerArgs &a, hipStream_t s);
hipError_t launch_smt_v1(double *data, int SX, int SY, int nz, int sx, int sy, double diag, hipStream_t s);
// the singular column of several solves (batched form): one block per item
struct SmtItem {
    long offset;  // element offset of the solve's field in the data arena
    int SX, SY, nz, sx, sy;
    double diag;
};
hipError_t launch_smt_v1_batch(double *data, const SmtItem *items, int nitems, int nz_max, hipStream_t s);

// counter-generator right-hand side (pp_solver.hip): out[i] = 2 u - 1,
// u = (splitmix64(seed * 2^32 + i) >> 11) * 2^-53, optionally rounded to FP32
hipError_t launch_fill_uniform(double *out, long n, unsigned long long seed, int round_f32, hipStream_t s);

}  // namespace pp
