// pp_mixed.h -- shared declarations of the mixed-precision kernels (pp_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "pp_zsolve.h"

namespace pp {

// arguments of the FP64 residual r = f_eff - L u
struct ResArgs {
    const double *f;  // right-hand side as given (unscaled)
    const double *u;  // current solution, FP64 (the residual launch's source kind RES_U64)
    // or its FP32 parts: u0 (RES_U0) or u0 + u1 (RES_U0_U1), summed in FP64
    const float *u0, *u1;
    void *r;          // residual out (FP32)
    int nx, ny, nz;
    double c3;        // dz: f_eff = f*dz*dz - Dirichlet term
    double Cx, Cy;    // (dz/dx)^2, (dz/dy)^2
    char bc[6];
    const double *dtab;  // Dirichlet boundary tables (x, y, z profiles, face values)
    int dmask;           // bit f set: face f carries a Dirichlet term
    // pinned (0,0) column: the pinned system replaces row 0 of the singular z operator, so
    // the solution leaves a plane-constant residual d0/(NxNy) at z = 0 that belongs to no
    // other mode; it is removed here (null: no pinned column)
    const double *pin_defect;
    // pinned (0,0) column: per-block partial sums of u over plane 0, or over planes 0 and
    // nz-1 for periodic z (null: not wanted); residual_parts(nx, ny) entries
    double *upart;
};

// where the residual kernel reads the current solution from
enum ResSource : int { RES_U64 = 0, RES_U0 = 1, RES_U0_U1 = 2 };

// per-block partial sums of u over plane 0 written by the residual kernel
int residual_parts(int nx, int ny);
// blocks of one residual launch over an nx x ny x nz grid (residual_tile_x() x
// residual_tile_y() cells, residual_zplanes() planes per block)
int residual_blocks(int nx, int ny, int nz);
int residual_zplanes();
int residual_tile_x();
int residual_tile_y();

// one pinned member of a batched (0,0)-column launch
struct Mode00Item {
    const double *line_sums;
    int ny, nz;
    double b0, bn, inv_scale;
    double *plane_add;
    const double *upart;
    int nparts;
    int zper;  // periodic z: the reference's singular cyclic column (row 0 averaged)
};

// cols[i] = {sx, sy, member, 0}; nz_max bounds every listed column's length
hipError_t launch_zfix64_f32(const ZArgs &a, const int4 *cols, int ncols, int nz_max, hipStream_t s);
hipError_t launch_mode00_batch(const Mode00Item *items, int nitems, int nz_max, int correct, hipStream_t s);
// two: some member has nx = 1 (mod residual_tile_x()) or ny = 1 (mod residual_tile_y())
// (residual_two_halo), which takes a second halo load per direction
bool residual_two_halo(int nx, int ny);
hipError_t launch_residual_f32_batch(const ResArgs *solves, const int4 *items, int nitems, bool two, int src,
                                     hipStream_t s);
// grids with even nx take the two-cells-per-lane kernel (residual_tile_x_of(nx) = 128 columns
// per tile, partial sums in the same layout); its batched launch: items {member, bx, by, bz}
// over 128-column tiles, twoy: some member has ny = 1 (mod residual_tile_y())
bool residual_pairs(int nx);
int residual_tile_x_of(int nx);
hipError_t launch_residual_pair_f32_batch(const ResArgs *solves, const int4 *items, int nitems, bool twoy, int src,
                                          hipStream_t s);
hipError_t launch_mode00(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                         double *plane_add, const double *upart, int nparts, int correct, int zper, hipStream_t s);
// src: ResSource (which of a.u / a.u0 / a.u1 hold the current solution)
hipError_t launch_residual_f32(const ResArgs &a, int src, hipStream_t s);

}  // namespace pp
