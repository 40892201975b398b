// pp_mixed.h -- shared declarations of the mixed-precision kernels (pp_mixed.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "pp_zsolve.h"

namespace pp {

// arguments of the FP64 residual r = f_eff - L u
struct ResArgs {
    const double *f;  // right-hand side as given (unscaled)
    const double *u;  // current solution
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
    // pinned (0,0) column: per-block partial sums of u over plane 0 (null: not wanted);
    // residual_parts(nx, ny) entries
    double *upart;
};

int residual_parts(int nx, int ny);

hipError_t launch_zfix64_f32(const ZArgs &a, const int2 *cols, int ncols, hipStream_t s);
hipError_t launch_mode00(const double *line_sums, int ny, int nz, double b0, double bn, double inv_scale,
                         double *plane_add, const double *upart, int nparts, int correct, hipStream_t s);
hipError_t launch_residual_f32(const ResArgs &a, hipStream_t s);

}  // namespace pp
