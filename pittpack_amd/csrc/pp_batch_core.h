// pp_batch_core.h -- the engine of the batched form, shared by pp_batch (one batch, its own
// device memory) and pp_server (many batches in flight, memory from one pool).
//
// A batch is planned on the host (member offsets, per-member descriptors, per-launch block
// items), then bound to one contiguous device region, uploaded asynchronously and run on a
// stream.  Every stage is one launch per shape class: line length and transform kind for
// the x and y transforms, rows per lane for the z solve.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "pp_launch.h"
#include "pp_line.h"
#include "pp_mixed.h"
#include "pp_setup.h"
#include "pp_zsolve.h"

namespace pp {

struct AxisDev {
    double *tw = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr;
};

// FP64 twiddle tables per line length, resident on the device until destruction
class TableCache {
  public:
    ~TableCache();
    int get(int N, AxisDev &out);

  private:
    std::map<int, AxisDev> t_;
};

// events recorded between the launches of one run (null ev: no timing)
struct Marks {
    hipEvent_t *ev = nullptr;
    int n = 0, max = 0;
};

struct BatchMember {
    SolveSetup g;
    long offset = 0;  // element offset of the member's field in the rhs / solution / work arenas
    long lam_off[2] = {0, 0};
    long dtab_off = 0;
    bool pinned = false;  // mixed form: pinned (0,0) column solved in FP64 from plane sums
    long sums_off = 0, plane_off = 0, upart_off = 0;
    long toff = 0;  // FP64: offset of the member's z tiles in the tile arena (-1: none)
};

class BatchCore {
  public:
    // options (set before plan)
    int precision = 0;  // 0 FP64, 1 FP32 + FP64 refinement
    int refine = 2;
    int zero_mode = 1;
    double tau = 0.05;
    double zthomas_tau = 1e-3;  // FP64: columns below it by the reference's Thomas sweeps (0: off)
    int force_var = -1;  // FP64 line classes: -1 tuned / default shapes, 0 default, 1 alternative
    int fuse_res = 1;    // mixed form: corrections fused into the x forward where a member's rows allow

    std::vector<BatchMember> mem;

    // validates and appends a member; *id = its index
    int add(int nx, int ny, int nz, const char *bc, const double *X, int *id);
    // host plan: offsets, descriptors, launch items; sets region_bytes
    int plan();
    // device pointers inside [base, base + region_bytes), table pointers from tc
    int bind(char *base, TableCache *tc);
    // asynchronous uploads of the planned tables (host copies stay alive in this object)
    int upload(hipStream_t s);
    // the solve of every member (right-hand sides kept, solutions in the solution arena)
    int run(hipStream_t s, Marks *mk);
    // FP64: time each line class in both launch shapes on this batch (bound, uploaded) and
    // keep the faster per class and direction; the right-hand-side arena is overwritten
    int tune_lines(hipStream_t s);
    int alt_classes() const;

    long long points() const;
    size_t region_bytes = 0;
    bool planned = false, bound = false;

    // device pointers (valid after bind)
    double *d_rhs = nullptr, *d_sol = nullptr;
    double *d_tiles = nullptr;  // FP64: z tiles between the y transforms and the z solve
    float *d_w = nullptr;
    float *d_u0 = nullptr, *d_cacc = nullptr;  // mixed form: the solution's FP32 parts while refined
    long arena_elems = 0;
    int launches[3] = {0, 0, 0};

  private:
    struct HostClass {
        int key = 0, kind = 0, nitems = 0;
        std::vector<LineItem> li;
        std::vector<ZItem> zi;
        size_t off = 0;
        void *d_items = nullptr;
        // FP64 line classes: launch shape per direction (forward, inverse) and, for x, the
        // block items of the alternative shape (its lines per block differ)
        int var[2] = {0, 0};
        std::vector<LineItem> li_alt;
        size_t off_alt = 0;
        void *d_items_alt = nullptr;
        bool fused = false;  // correction x classes: members whose residual runs in k_resx
    };
    int launch_class(hipStream_t s, const HostClass &c, int form, int axis, int fwd, const void *in, void *out,
                     int bflags, double *sums, const double *plane, int accumulate, int ztiled,
                     const float *acc0 = nullptr, const float *acc1 = nullptr);
    template <typename R>
    int plan_lines(int axis, std::vector<HostClass> &out);
    void plan_z();
    int plan_mixed();
    int run_line_stage(hipStream_t s, std::vector<HostClass> &cls, int form, int axis, int fwd, const void *in,
                       void *out, int bflags, double *sums, const double *plane, int accumulate, int ztiled = 0,
                       const float *acc0 = nullptr, const float *acc1 = nullptr);
    int run_z_stage(hipStream_t s, bool f32);
    int run_fp64(hipStream_t s, Marks *mk);
    int mixed_inner(hipStream_t s, Marks *mk);
    int run_mixed(hipStream_t s, Marks *mk);

    // host copies of everything uploaded
    std::vector<double> lam_, dtab_;
    std::vector<LineSolve> ls_;
    std::vector<ZSolve> zs_;
    std::vector<HostClass> lx_, ly_, lz_;
    // mixed form: the corrections' x forward, each x class split into the members whose
    // residual is computed inside it (k_resx: ny a multiple of its lines per block, as the
    // single-solve handle decides) and the rest (residual kernel, then the FP32 transform)
    std::vector<HostClass> lxc_;
    std::vector<char> mfused_;
    // periodic z: one-pass classes (key nz, a compiled complex plan; items {member, column
    // block, y row}), members of other lengths (real z-line transforms, one launch each), the
    // singular columns of FP64 members, and 2 cos(2 pi k / nz) per distinct periodic length
    std::vector<HostClass> lzp_;
    std::vector<int> zpl_;
    std::vector<SmtItem> smt_;
    std::vector<double> zc2_;
    std::map<int, long> zc2_off_;
    // FP64 members with 513..1024-row columns: the long-column kernel, one launch each
    std::vector<int> zlong_;
    // FP64: nearly singular columns of the PCR classes' members, solved by the reference's
    // Thomas sweeps after the z classes, and their factor tables
    std::vector<ZThomasItem> zth_;
    std::vector<double> zth_tab_;
    std::vector<int4> fix_;
    std::vector<HostClass> lfix_;  // 16-column tiles holding near-singular columns, per J
    std::vector<Mode00Item> m00_;
    std::vector<int> m00_mem_;  // member of each m00_ entry
    std::vector<ResArgs> res_;
    std::vector<int4> ritems_;   // residual tiles of odd-nx members (64 columns)
    std::vector<int4> rpitems_;  // residual tiles of even-nx members (two cells per lane, 128 columns)
    bool res_two_ = false;   // some odd-nx member's residual tiles need the second halo load
    bool resp_two_ = false;  // some even-nx member has ny = 1 (mod 4)
    int nz_max_ = 0;
    long nsums_ = 0, nplane_ = 0, nupart_ = 0, ntiles_ = 0;
    // region layout (byte offsets)
    size_t o_zth_ = 0, o_zth_tab_ = 0, o_u0_ = 0, o_cacc_ = 0, o_zc2_ = 0, o_smt_ = 0, o_rhs_ = 0, o_sol_ = 0, o_tiles_ = 0, o_w_ = 0, o_lam_ = 0, o_dtab_ = 0, o_ls_ = 0, o_zs_ = 0, o_sums_ = 0, o_plane_ = 0,
           o_upart_ = 0, o_fix_ = 0, o_m00_ = 0, o_res_ = 0, o_ritems_ = 0, o_rpitems_ = 0;
    std::map<int, AxisDev> tabs_;
    // device pointers of the tables
    double *d_lam_ = nullptr, *d_dtab_ = nullptr, *d_sums_ = nullptr, *d_plane_ = nullptr, *d_upart_ = nullptr;
    LineSolve *d_ls_ = nullptr;
    ZSolve *d_zs_ = nullptr;
    int4 *d_fix_ = nullptr, *d_ritems_ = nullptr, *d_rpitems_ = nullptr;
    double *d_zc2_ = nullptr;
    ZThomasItem *d_zth_ = nullptr;
    double *d_zth_tab_ = nullptr;
    SmtItem *d_smt_ = nullptr;
    int smt_nz_max_ = 0;

    Mode00Item *d_m00_ = nullptr;
    ResArgs *d_res_ = nullptr;
};

}  // namespace pp
