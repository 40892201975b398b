// pp_handle.h -- state of a single-solve handle (pp_handle of include/pittpack_hip.h), shared by
// the translation units that drive its FP64 pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "pp_launch.h"
#include "pp_setup.h"
#include "pp_zsolve.h"

namespace pp {

// device table pointers of one axis (twiddles and transform tables, pp_setup.h)
struct DevTables {
    void *tw = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr;
};

}  // namespace pp

struct pp_solver {
    pp::SolveSetup g;  // sizes, boundary conditions, box, spacing, eigenvalue parameters
    bool bc_set = false, box_set = false, coords_set = false, rhs_set = false;
    int zero_mode = 1;
    int zsolver = 0;
    bool timing = false;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double *d_field = nullptr;
    double *d_lamx = nullptr, *d_lamy = nullptr, *d_dtab = nullptr;
    double *d_ztw = nullptr, *d_zc2 = nullptr;  // periodic z: twiddles and 2 cos(2 pi k / nz)
    double *d_zt = nullptr;                     // z tiles of the FP64 form
    int ztiles = 1;                             // PP_OPT_ZLAYOUT
    // autotuner (PP_OPT_AUTOTUNE): launch shape per FP64 line stage (x fwd, y fwd, y inv,
    // x inv; 1 = line_threads_alt) and the z configuration, chosen by timing
    int autotune = 0;
    bool tuned = false;
    int line_var[4] = {0, 0, 0, 0};
    int force_var = -1;                         // PP_OPT_LINE_VARIANT (every stage; -1: tuner's)
    // PP_OPT_GRAPH: repeated solves on the same device pointers replay a HIP graph of the
    // solve's launches (captured on the second such solve; any configuration change drops it)
    int graph_mode = 0;
    hipGraphExec_t graph = nullptr;
    const void *g_rhs = nullptr;
    const void *g_out = nullptr;
    int g_eager = 0;
    hipEvent_t *ev_force = nullptr;             // tuner's own events for the launch marks
    int mg_cycles = 200;                        // multigrid z solver: V-cycle limit
    double mg_tol = 1e-16;                      //   and relative max-norm residual tolerance
    // FP64 form: columns with |eig+2| below zthomas_tau (not pinned) are solved by the
    // reference's Thomas sweeps after the PCR kernel (0: off); their list, built with the tables
    double zthomas_tau = 1e-3;
    pp::ZThomasItem *d_zthomas = nullptr;
    double *d_zthomas_tab = nullptr;  // their Thomas factors (zthomas_factors), 2 nz per column
    int nzthomas = 0;
    bool zthomas_ready = false;
    int zsing_x = -1, zsing_y = -1;             // periodic z: the singular column, if any
    double zsing_eig = -2.0;
    std::vector<double> lamx_h, lamy_h;  // host copies of the eigenvalue factors
    pp::DevTables tab64[2];
    pp::DevTables tabz;  // periodic z without a compiled complex plan: real z-line tables
    // kernel-timing ring: slot s holds up to kMaxMarks events recorded between the kernel
    // launches of one solve
    static constexpr int kMaxMarks = 64;
    std::vector<hipEvent_t> ev;
    int ev_slots = 0, ev_next = 0, ev_used = 0;
    hipEvent_t *ev_cur = nullptr;
    int marks_cur = 0, marks_last = 0;
    bool tables_ready = false;
    // mixed precision form
    int precision = 0;   // 0 FP64, 1 FP32 + FP64 refinement
    int refine = 2;
    double tau = 0.05;
    float *d_w = nullptr;          // FP32 work field
    float *d_u0 = nullptr;         // FP32 initial solution of the refinement (u = u0 + c1 + ... + cR)
    float *d_cacc = nullptr;       // FP32 sum of the corrections before the last
    float *d_wt = nullptr;         // FP32 z tiles (32 columns x 16 J rows) between the y transforms
    int wtiles = 1;                // mixed form: hand the z solve FP32 z tiles where the sizes allow
    int wvar = 0;                  // mixed form, FP32 tiled y stages: 1 = half the threads per column
    int fuse_res = 1;              // PP_OPT_FUSE_RESIDUAL: correction's residual inside its x forward
    bool mtuned = false;           // mixed form tuned (PP_OPT_AUTOTUNE)
    double *d_aux = nullptr;       // FP64 scratch field (RHS copy or solution) for host-facing calls
    double *d_line_sums = nullptr; // ny*nz
    double *d_plane_add = nullptr; // nz + 2
    double *d_upart = nullptr;     // residual kernel's plane-0 partial sums of u
    int4 *d_fix = nullptr;         // near-singular columns solved in FP64: {sx, sy, 0, 0}
    int nfix = 0;
    pp::ZItem *d_fixtiles = nullptr;   // the 16-column tiles holding them
    int nfixtiles = 0;
    pp::ZSolve *d_zself = nullptr;     // this handle as a one-member batch (tile items read it)
    bool pinned00 = false;
    bool mixed_ready = false;
    int io_format = 0;             // PP_OPT_IO_FORMAT
    // manufactured solution of the reference (pp_init_trigonometric): -1 none (pp_get_error
    // then uses the cos(2 pi z) case), 0 / 1 real / imaginary part of its complex field
    int trig_part = -1;
    double *d_partials = nullptr;  // pp_get_error's per-block partial sums
};


namespace pp {
// the five FP64 stages of a handle whose tables are in place (tables_ready): rhs -> field
int run_fp64(pp_solver *h, const double *rhs, double *field);
}  // namespace pp
