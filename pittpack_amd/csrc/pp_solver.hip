// pp_solver.hip -- host side of the MI355X PittPack solve path: handle state, boundary
// condition set-up, device tables, kernel dispatch and the C ABI of include/pittpack_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "pittpack_hip.h"
#include "pp_line.h"
#include "pp_line_bs.h"
#include "pp_launch.h"
#include "pp_mixed.h"
#include "pp_resx.h"
#include "pp_setup.h"
#include "pp_zsolve.h"
#include "pp_handle.h"

namespace pp {
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

using namespace pp;

namespace pp {
extern template hipError_t launch_line_kind<KIND_DD>(int, int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_kind<KIND_NN>(int, int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_kind<KIND_PP>(int, int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_kind<KIND_ND>(int, int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_kind<KIND_DN>(int, int, int, int, const LineArgs &, hipStream_t);

extern template hipError_t launch_resx_kind<KIND_DD>(int, int, const LineArgs &, const ResArgs &, const ResArgs *, hipStream_t);
extern template hipError_t launch_resx_kind<KIND_NN>(int, int, const LineArgs &, const ResArgs &, const ResArgs *, hipStream_t);
extern template hipError_t launch_resx_kind<KIND_PP>(int, int, const LineArgs &, const ResArgs &, const ResArgs *, hipStream_t);
extern template hipError_t launch_resx_kind<KIND_ND>(int, int, const LineArgs &, const ResArgs &, const ResArgs *, hipStream_t);
extern template hipError_t launch_resx_kind<KIND_DN>(int, int, const LineArgs &, const ResArgs &, const ResArgs *, hipStream_t);

extern template hipError_t launch_line_bs_kind<KIND_DD>(int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_bs_kind<KIND_NN>(int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_bs_kind<KIND_PP>(int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_bs_kind<KIND_ND>(int, int, int, const LineArgs &, hipStream_t);
extern template hipError_t launch_line_bs_kind<KIND_DN>(int, int, int, const LineArgs &, hipStream_t);

// line transforms: one translation unit per kind (pp_line_<kind>.hip for the compiled
// lengths, pp_line_bs_<kind>.hip for the others)
hipError_t launch_line(int form, int kind, int N, int AX, int FWD, const LineArgs &a0, hipStream_t s)
{
    if (!line_compiled(N)) {
        LineArgs a = a0;
        a.n = N;
        switch (kind) {
        case KIND_DD: return launch_line_bs_kind<KIND_DD>(form, AX, FWD, a, s);
        case KIND_NN: return launch_line_bs_kind<KIND_NN>(form, AX, FWD, a, s);
        case KIND_PP: return launch_line_bs_kind<KIND_PP>(form, AX, FWD, a, s);
        case KIND_ND: return launch_line_bs_kind<KIND_ND>(form, AX, FWD, a, s);
        case KIND_DN: return launch_line_bs_kind<KIND_DN>(form, AX, FWD, a, s);
        default: return hipErrorInvalidValue;
        }
    }
    const LineArgs &a = a0;
    switch (kind) {
    case KIND_DD: return launch_line_kind<KIND_DD>(form, N, AX, FWD, a, s);
    case KIND_NN: return launch_line_kind<KIND_NN>(form, N, AX, FWD, a, s);
    case KIND_PP: return launch_line_kind<KIND_PP>(form, N, AX, FWD, a, s);
    case KIND_ND: return launch_line_kind<KIND_ND>(form, N, AX, FWD, a, s);
    case KIND_DN: return launch_line_kind<KIND_DN>(form, N, AX, FWD, a, s);
    default: return hipErrorInvalidValue;
    }
}

// the mixed form's fused correction step (pp_resx.h): residual of the current solution (src:
// RES_U0 or RES_U0_U1) into the FP32 x-forward transform of kind `kind`
hipError_t launch_resx(int kind, int N, int src, const LineArgs &a, const ResArgs &r, const ResArgs *rsolves,
                       hipStream_t s)
{
    switch (kind) {
    case KIND_DD: return launch_resx_kind<KIND_DD>(N, src, a, r, rsolves, s);
    case KIND_NN: return launch_resx_kind<KIND_NN>(N, src, a, r, rsolves, s);
    case KIND_PP: return launch_resx_kind<KIND_PP>(N, src, a, r, rsolves, s);
    case KIND_ND: return launch_resx_kind<KIND_ND>(N, src, a, r, rsolves, s);
    case KIND_DN: return launch_resx_kind<KIND_DN>(N, src, a, r, rsolves, s);
    default: return hipErrorInvalidValue;
    }
}

__global__ void k_fill_uniform(double *out, long n, unsigned long long seed, int round_f32)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = (seed << 32) + (unsigned long long)i + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    double v = 2.0 * u - 1.0;
    if (round_f32) v = (double)(float)v;
    out[i] = v;
}

hipError_t launch_fill_uniform(double *out, long n, unsigned long long seed, int round_f32, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, seed, round_f32);
    return hipGetLastError();
}

}  // namespace pp
namespace {

bool size_supported(int n) { return line_size_supported(n); }

}  // namespace

namespace {

#define PP_HIP_CHECK(call)                       \
    do {                                         \
        hipError_t e_ = (call);                  \
        if (e_ != hipSuccess) return PP_HIP_ERROR; \
    } while (0)

int mark(pp_solver *h)
{
    if (h->ev_cur && h->marks_cur < pp_solver::kMaxMarks) {
        PP_HIP_CHECK(hipEventRecord(h->ev_cur[h->marks_cur++], h->stream));
    }
    return PP_SUCCESS;
}

int begin_timing(pp_solver *h)
{
    h->ev_cur = nullptr;
    if (h->ev_force) {
        h->ev_cur = h->ev_force;
        h->marks_cur = 0;
        return mark(h);
    }
    if (!h->timing) return PP_SUCCESS;
    h->ev_cur = &h->ev[(size_t)pp_solver::kMaxMarks * h->ev_next];
    h->ev_next = (h->ev_next + 1) % h->ev_slots;
    if (h->ev_used < h->ev_slots) h->ev_used++;
    h->marks_cur = 0;
    return mark(h);
}

void end_timing(pp_solver *h)
{
    if (h->ev_cur && !h->ev_force) h->marks_last = h->marks_cur;
    h->ev_cur = nullptr;
}

// drop the captured solve (after the stream has drained: a replay may still be running)
void drop_graph(pp_solver *h)
{
    if (h->graph) {
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(h->graph);
    }
    h->graph = nullptr;
    h->g_rhs = h->g_out = nullptr;
    h->g_eager = 0;
}

// launch shape of FP64 line stage i (0 x fwd, 1 y fwd, 2 y inv, 3 x inv) of length N
int stage_var(const pp_solver *h, int i, int N)
{
    const int v = h->force_var >= 0 ? h->force_var : h->line_var[i];
    return v && line_has_variant(N) ? 1 : 0;
}

void free_mixed(pp_solver *h)
{
    if (h->d_w) (void)hipFree(h->d_w);
    if (h->d_wt) (void)hipFree(h->d_wt);
    if (h->d_u0) (void)hipFree(h->d_u0);
    if (h->d_cacc) (void)hipFree(h->d_cacc);
    h->d_wt = nullptr;
    h->d_u0 = h->d_cacc = nullptr;
    if (h->d_aux) (void)hipFree(h->d_aux);
    if (h->d_line_sums) (void)hipFree(h->d_line_sums);
    if (h->d_plane_add) (void)hipFree(h->d_plane_add);
    if (h->d_upart) (void)hipFree(h->d_upart);
    if (h->d_fix) (void)hipFree(h->d_fix);
    if (h->d_fixtiles) (void)hipFree(h->d_fixtiles);
    if (h->d_zself) (void)hipFree(h->d_zself);
    h->d_fixtiles = nullptr;
    h->d_zself = nullptr;
    h->nfixtiles = 0;
    h->d_w = nullptr;
    h->d_aux = nullptr;
    h->d_line_sums = h->d_plane_add = h->d_upart = nullptr;
    h->d_fix = nullptr;
    h->nfix = 0;
    h->mixed_ready = false;
}

void free_zthomas(pp_solver *h)
{
    if (h->d_zthomas) (void)hipFree(h->d_zthomas);
    if (h->d_zthomas_tab) (void)hipFree(h->d_zthomas_tab);
    h->d_zthomas = nullptr;
    h->d_zthomas_tab = nullptr;
    h->nzthomas = 0;
    h->zthomas_ready = false;
}

void free_tables(pp_solver *h)
{
    free_zthomas(h);
    for (DevTables *t : {&h->tab64[0], &h->tab64[1], &h->tabz}) {
        for (void **p : {&t->tw, &t->A, &t->B, &t->C, &t->D}) {
            if (*p) (void)hipFree(*p);
            *p = nullptr;
        }
    }
    for (double **p : {&h->d_lamx, &h->d_lamy, &h->d_dtab, &h->d_ztw, &h->d_zc2}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    h->tables_ready = false;
}

int upload(void **dst, const std::vector<double> &v)
{
    if (hipMalloc(dst, v.size() * sizeof(double)) != hipSuccess) return PP_ALLOCATION_FAIL;
    if (hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return PP_HIP_ERROR;
    return PP_SUCCESS;
}

// Build everything that depends on (sizes, bc, box): eigenvalue factors, twiddles and the
// Dirichlet boundary term tables.  Mirrors assignBoundary/setEigenVal/setScale/setCoords.
int build_tables(pp_solver *h)
{
    free_tables(h);
    const int n[2] = {h->g.nx, h->g.ny};
    int rc;
    for (int a = 0; a < 2; a++) {
        const AxisTables &t = axis_tables_cached(n[a]);
        DevTables &d = h->tab64[a];
        if ((rc = upload(&d.tw, t.tw))) return rc;
        if ((rc = upload(&d.A, t.A))) return rc;
        if ((rc = upload(&d.B, t.B))) return rc;
        if ((rc = upload(&d.C, t.C))) return rc;
        if ((rc = upload(&d.D, t.D))) return rc;
    }
    const std::vector<double> lx = eig_factors(h->g, 0), ly = eig_factors(h->g, 1);
    if ((rc = upload((void **)&h->d_lamx, lx))) return rc;
    if ((rc = upload((void **)&h->d_lamy, ly))) return rc;
    h->lamx_h = lx;
    h->lamy_h = ly;
    if ((rc = upload((void **)&h->d_dtab, dirichlet_table(h->g)))) return rc;
    h->zsing_x = h->zsing_y = -1;
    if (h->g.zper) {
        AxisTables zt;
        if (zperiodic_supported(h->g.nz)) {
            make_axis_tables(2 * h->g.nz, zt);  // complex FFT of length nz
            if ((rc = upload((void **)&h->d_ztw, zt.tw))) return rc;
        } else {
            // real z-line transforms of length nz (R2HC / HC2R)
            make_axis_tables(h->g.nz, zt);
            DevTables &d = h->tabz;
            if ((rc = upload(&d.tw, zt.tw)) || (rc = upload(&d.A, zt.A)) || (rc = upload(&d.B, zt.B)) ||
                (rc = upload(&d.C, zt.C)) || (rc = upload(&d.D, zt.D)))
                return rc;
        }
        if ((rc = upload((void **)&h->d_zc2, zper_cos2(h->g.nz)))) return rc;
        const double Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
        const double Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
        for (int sy = 0; sy < h->g.ny && h->zsing_x < 0; sy++)
            for (int sx = 0; sx < h->g.nx; sx++)
                if (std::fabs(column_eig(lx, ly, sx, sy, Cx, Cy) + 2.) < 1.e-10) {
                    h->zsing_x = sx;
                    h->zsing_y = sy;
                    h->zsing_eig = column_eig(lx, ly, sx, sy, Cx, Cy);
                    break;
                }
    }
    h->tables_ready = true;
    return PP_SUCCESS;
}

// z lengths each FP64 z solver takes (non-periodic z)
bool fp64_z_supported(int zsolver, int nz)
{
    switch (zsolver) {
    case 1: return zsolve_supported_f64(nz);  // partitioned Thomas: up to 32 rows per lane
    case 2: return nz >= 4 && 6 * (size_t)nz * sizeof(double) <= 64 * 1024;  // multigrid: 6 LDS rows
    case 3:
    case 4: return zsolve_alt_supported(nz);
    default: return zsolve_supported_f64(nz) || zsolve_long_supported_f64(nz);
    }
}

int ensure_ready(pp_solver *h)
{
    if (!h->bc_set || !h->coords_set) return PP_CALL_ORDER;
    if (h->g.zper) {
        // periodic z: any z length from 2 to 1024 (the mixed form's singular column: from 3)
        if (!zperiodic_supported(h->g.nz) && !line_size_supported(h->g.nz)) return PP_UNSUPPORTED_SIZE;
        if (h->precision == 1 && h->g.nz < 3) return PP_UNSUPPORTED_SIZE;
    }
    if (h->precision == 1 && !h->g.zper && !zsolve_supported_f32(h->g.nz)) return PP_UNSUPPORTED_SIZE;
    if (h->precision == 0 && !h->g.zper && !fp64_z_supported(h->zsolver, h->g.nz)) return PP_UNSUPPORTED_SIZE;
    if (!h->tables_ready) {
        h->mixed_ready = false;
        return build_tables(h);
    }
    return PP_SUCCESS;
}

LineArgs line_args(pp_solver *h, int axis, double *field)
{
    LineArgs a{};
    a.in = field;
    a.out = field;
    const DevTables &t = h->tab64[axis];
    a.tw = (const double *)t.tw;
    a.A = (const double *)t.A;
    a.B = (const double *)t.B;
    a.C = (const double *)t.C;
    a.D = (const double *)t.D;
    a.in_scale = 1.0;
    a.out_scale = 1.0;
    a.dtab = h->d_dtab;
    a.dmask = 0;
    a.nx = h->g.nx;
    a.ny = h->g.ny;
    a.nz = h->g.nz;
    if (axis == 0) {
        a.nlines = (long)h->g.ny * h->g.nz;
    } else {
        a.nlines = h->g.nx;
        a.ncols = h->g.nx;
        a.nplanes = h->g.nz;
    }
    return a;
}

// the five stages, FP64: rhs -> field (x forward), then in place on field
// z-tiled layout between the y transforms and the z solve: the y forward writes each
// (x/16, y) tile of nz rows x 16 columns (one 128-byte line per row) contiguously, so the z
// solve reads and writes whole tiles instead of one row per z plane.  y blocks of 16 columns
// fill a tile, y blocks of 8 (1280- to 2048-byte y lines' LDS slices) half of one; z blocks
// take a whole tile (up to 512 rows) or half (longer columns).  z not periodic.
bool use_ztiles(pp_solver *h)
{
    // (the partitioned solver rounds its rows per lane up to a coarser set, so it keeps the
    // in-place layout)
    if (!h->ztiles || h->zsolver != 0 || h->g.zper) return false;
    if (!zsolve_supported_f64(h->g.nz) && !zsolve_long_supported_f64(h->g.nz)) return false;
    const int ly = line_lines_any<double>(h->g.ny, 1);
    return ly > 0 && zsolve_cpb<double>() % ly == 0;
}

// rows per z tile: whole lane rows of the z solve (16 J), zero padding below nz
int ztile_rows(int nz) { return (nz + 15) / 16 * 16; }

int ensure_ztiles(pp_solver *h)
{
    if (h->d_zt) return PP_SUCCESS;
    const size_t W = zsolve_cpb<double>();  // the tile width
    const size_t elems = ((size_t)h->g.nx + W - 1) / W * W * h->g.ny * ztile_rows(h->g.nz);
    if (hipMalloc((void **)&h->d_zt, elems * sizeof(double)) != hipSuccess) {
        h->d_zt = nullptr;
        return PP_ALLOCATION_FAIL;
    }
    // the padding (rows from nz, columns from nx) is never written by the transforms; zero
    // keeps it finite through the z solve, which treats those rows as identity rows
    PP_HIP_CHECK(hipMemset(h->d_zt, 0, elems * sizeof(double)));
    return PP_SUCCESS;
}

// Periodic z solve of the field in place: the one-pass kernel for lengths with a compiled
// complex plan, else real z-line transforms (the columns of a [nz][ny*nx] array: R2HC,
// divide, HC2R).  f32: the mixed form's FP32 work field, singular column written as zeros.
int run_zperiodic(pp_solver *h, void *data, bool f32)
{
    hipStream_t s = h->stream;
    ZPerArgs zp{};
    zp.data = data;
    zp.SX = h->g.nx;
    zp.SY = h->g.ny;
    zp.nz = h->g.nz;
    zp.lamx = h->d_lamx;
    zp.lamy = h->d_lamy;
    zp.Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
    zp.Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
    zp.tw = h->d_ztw;
    zp.c2 = h->d_zc2;
    zp.pinned_zero = f32 ? 1 : 0;
    if (zperiodic_supported(h->g.nz)) {
        PP_HIP_CHECK(f32 ? launch_zperiodic_f32(zp, s) : launch_zperiodic_f64(zp, s));
        return PP_SUCCESS;
    }
    LineArgs a{};
    a.in = data;
    a.out = data;
    const DevTables &t = h->tabz;
    a.tw = (const double *)t.tw;
    a.A = (const double *)t.A;
    a.B = (const double *)t.B;
    a.C = (const double *)t.C;
    a.D = (const double *)t.D;
    a.in_scale = 1.0;
    a.out_scale = 1.0;
    a.nlines = (long)h->g.nx * h->g.ny;
    a.ncols = h->g.nx * h->g.ny;
    a.nplanes = 1;
    const int form = f32 ? FORM_F32 : FORM_F64;
    PP_HIP_CHECK(launch_line(form, KIND_PP, h->g.nz, 1, 1, a, s));
    PP_HIP_CHECK(f32 ? launch_zper_divide_f32(zp, s) : launch_zper_divide(zp, s));
    PP_HIP_CHECK(launch_line(form, KIND_PP, h->g.nz, 1, 0, a, s));
    return PP_SUCCESS;
}

// PP_OPT_ZTHOMAS_TAU applies: FP64 PCR -> Thomas or partitioned Thomas on columns of at
// most 512 rows, non-periodic z
bool zthomas_on(const pp_solver *h)
{
    return h->zthomas_tau > 0.0 && !h->g.zper && (h->zsolver == 0 || h->zsolver == 1) && zsolve_supported_f64(h->g.nz);
}

// The columns the serial Thomas takes: |eig+2| < zthomas_tau with the kernels' eigenvalue
// rounding, the pinned column excluded (it keeps the row-0 rule); built once per tables,
// threshold and zero-mode rule
int ensure_zthomas(pp_solver *h)
{
    if (h->zthomas_ready) return PP_SUCCESS;
    // the host copies of the eigenvalue factors come with the tables
    if ((int)h->lamx_h.size() != h->g.nx || (int)h->lamy_h.size() != h->g.ny) return PP_CALL_ORDER;
    if (h->d_zthomas || h->d_zthomas_tab) {
        PP_HIP_CHECK(hipStreamSynchronize(h->stream));
        if (h->d_zthomas) PP_HIP_CHECK(hipFree(h->d_zthomas));
        if (h->d_zthomas_tab) PP_HIP_CHECK(hipFree(h->d_zthomas_tab));
        h->d_zthomas = nullptr;
        h->d_zthomas_tab = nullptr;
    }
    h->nzthomas = 0;
    int rc;
    const double Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
    const double Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
    std::vector<ZThomasItem> cols;
    std::vector<double> tab;
    const int n = h->g.nz;
    for (int sy = 0; sy < h->g.ny; sy++)
        for (int sx = 0; sx < h->g.nx; sx++) {
            const double eig = column_eig(h->lamx_h, h->lamy_h, sx, sy, Cx, Cy);
            bool pin = std::fabs(eig + 2.) < 1.e-10;
            if (pin && h->zero_mode == 1 && (h->g.zlo_d || h->g.zhi_d)) pin = false;
            if (pin || !(std::fabs(eig + 2.) < h->zthomas_tau)) continue;
            cols.push_back({sx, sy, -1, 0, (long)tab.size()});
            tab.resize(tab.size() + 2 * (size_t)n);
            zthomas_factors(eig, h->g.zlo_d ? eig - 1. : eig + 1., h->g.zhi_d ? eig - 1. : eig + 1., n,
                            tab.data() + tab.size() - 2 * (size_t)n);
        }
    if (!cols.empty()) {
        if (hipMalloc((void **)&h->d_zthomas, cols.size() * sizeof(ZThomasItem)) != hipSuccess) return PP_ALLOCATION_FAIL;
        PP_HIP_CHECK(hipMemcpy(h->d_zthomas, cols.data(), cols.size() * sizeof(ZThomasItem), hipMemcpyHostToDevice));
        if ((rc = upload((void **)&h->d_zthomas_tab, tab))) return rc;
    }
    h->nzthomas = (int)cols.size();
    h->zthomas_ready = true;
    return PP_SUCCESS;
}

// the five stages, FP64: rhs -> field (x forward), then on field (and the z tiles)
int run_fp64_stages(pp_solver *h, const double *rhs, double *field)
{
    hipStream_t s = h->stream;
    const bool zt = use_ztiles(h);
    int rc;
    if (zt && (rc = ensure_ztiles(h))) return rc;
    const bool zth = zthomas_on(h);
    if (zth && (rc = ensure_zthomas(h))) return rc;
    double *zdata = zt ? h->d_zt : field;
    if ((rc = begin_timing(h))) return rc;
    // 1. forward x: rhs * dz^2, Dirichlet boundary term, x transform
    {
        LineArgs a = line_args(h, 0, field);
        a.in = rhs;
        a.in_scale = h->g.dxyz[2];
        a.dmask = h->g.dmask;
        a.variant = stage_var(h, 0, h->g.nx);
        PP_HIP_CHECK(launch_line(FORM_F64, h->g.kind[0], h->g.nx, 0, 1, a, s));
    }
    if ((rc = mark(h))) return rc;
    // 2. forward y (into the z tiles)
    {
        LineArgs a = line_args(h, 1, field);
        a.out = zdata;
        a.ztiled = zt ? ZT_OUT : 0;
        a.variant = stage_var(h, 1, h->g.ny);
        a.tnz = ztile_rows(h->g.nz);
        a.tile_w = zt ? zsolve_cpb<double>() : 0;
        PP_HIP_CHECK(launch_line(FORM_F64, h->g.kind[1], h->g.ny, 1, 1, a, s));
    }
    if ((rc = mark(h))) return rc;
    // 3. z solve
    {
        ZArgs z{};
        z.data = zdata;
        z.tiled = zt ? 1 : 0;
        z.tile_w = zt ? zsolve_cpb<double>() : 0;
        z.tile_rows = zt ? ztile_rows(h->g.nz) : 0;
        z.SX = h->g.nx;
        z.SY = h->g.ny;
        z.nz = h->g.nz;
        z.lamx = h->d_lamx;
        z.lamy = h->d_lamy;
        z.Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
        z.Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
        z.zlo_dirichlet = h->g.zlo_d;
        z.zhi_dirichlet = h->g.zhi_d;
        z.zero_mode = h->zero_mode;
        if (h->g.zper) {
            if ((rc = run_zperiodic(h, field, false))) return rc;
            if (h->zsing_x >= 0)
                PP_HIP_CHECK(launch_smt_v1(field, h->g.nx, h->g.ny, h->g.nz, h->zsing_x, h->zsing_y, h->zsing_eig, s));
        } else {
            // nearly singular columns left to the reference's serial Thomas (PP_OPT_ZTHOMAS_TAU)
            if (zth) z.tau = h->zthomas_tau;
            switch (h->zsolver) {
            case 1: PP_HIP_CHECK(launch_zsolve_scan_f64(h->g.nz, z, s)); break;
            case 2: PP_HIP_CHECK(launch_zsolve_mg_f64(z, h->mg_cycles, h->mg_tol, s)); break;
            case 3: PP_HIP_CHECK(launch_zsolve_pcr_full_f64(z, s)); break;
            case 4: PP_HIP_CHECK(launch_zsolve_cr_f64(z, s)); break;
            default: PP_HIP_CHECK(launch_zsolve_pcr_f64(h->g.nz, z, s)); break;
            }
            if (zth) PP_HIP_CHECK(launch_zthomas_f64(z, h->d_zthomas, h->d_zthomas_tab, h->nzthomas, s));
        }
    }
    if ((rc = mark(h))) return rc;
    // 4. inverse y (from the z tiles)
    {
        LineArgs a = line_args(h, 1, field);
        a.in = zdata;
        a.ztiled = zt ? ZT_IN : 0;
        a.variant = stage_var(h, 2, h->g.ny);
        a.tnz = ztile_rows(h->g.nz);
        a.tile_w = zt ? zsolve_cpb<double>() : 0;
        PP_HIP_CHECK(launch_line(FORM_F64, h->g.kind[1], h->g.ny, 1, 0, a, s));
    }
    if ((rc = mark(h))) return rc;
    // 5. inverse x with the 1/(Nx*Ny) rescale fused
    {
        LineArgs a = line_args(h, 0, field);
        a.out_scale = 1. / (h->g.scale);
        a.variant = stage_var(h, 3, h->g.nx);
        PP_HIP_CHECK(launch_line(FORM_F64, h->g.kind[0], h->g.nx, 0, 0, a, s));
    }
    if ((rc = mark(h))) return rc;
    end_timing(h);
    return PP_SUCCESS;
}

// ---------------------------------------------------------------------------------------
// Autotuner of the FP64 form.  The choices, per grid shape and boundary kinds:
//   z configuration: PCR across 16 lanes then Thomas per lane on z tiles, the same in place,
//     or partitioned Thomas with cross-lane scans (where the PCR / Thomas split sits);
//   per line stage: the default launch shape or the alternative one (line_threads_alt:
//     radix plan and block shape).
// They are timed on the handle's shape with a scratch field (HIP events between the
// launches, best of a few solves), kept per process and reused by every handle of the same
// shape.  Results do not depend on the choice beyond rounding (each configuration is checked
// against the oracle by the tests).
// ---------------------------------------------------------------------------------------
struct TuneChoice {
    int zsolver, ztiles;
    int var[4];
};
std::mutex g_tune_mu;
std::map<std::array<int, 7>, TuneChoice> g_tune;

// best-of-`reps` time of each stage and of the whole solve on `buf` (in place)
int tune_time(pp_solver *h, double *buf, int reps, double stage[5], double &total)
{
    constexpr int K = 8;
    hipEvent_t ev[K];
    for (int i = 0; i < K; i++)
        if (hipEventCreate(&ev[i]) != hipSuccess) return PP_HIP_ERROR;
    for (int i = 0; i < 5; i++) stage[i] = 1e30;
    total = 1e30;
    int rc = run_fp64(h, buf, buf);  // warm-up (first-use set-up, caches)
    for (int r = 0; r < reps && !rc; r++) {
        h->ev_force = ev;
        rc = run_fp64(h, buf, buf);
        h->ev_force = nullptr;
        if (rc) break;
        if (hipEventSynchronize(ev[h->marks_cur - 1]) != hipSuccess) {
            rc = PP_HIP_ERROR;
            break;
        }
        for (int i = 0; i + 1 < h->marks_cur && i < 5; i++) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, ev[i], ev[i + 1]) != hipSuccess) rc = PP_HIP_ERROR;
            stage[i] = std::min(stage[i], (double)t);
        }
        float t = 0.f;
        if (hipEventElapsedTime(&t, ev[0], ev[h->marks_cur - 1]) != hipSuccess) rc = PP_HIP_ERROR;
        total = std::min(total, (double)t);
    }
    h->ev_force = nullptr;
    for (int i = 0; i < K; i++) (void)hipEventDestroy(ev[i]);
    return rc;
}

int autotune(pp_solver *h)
{
    const std::array<int, 7> key = {h->g.nx, h->g.ny, h->g.nz, h->g.kind[0], h->g.kind[1], h->g.zlo_d, h->g.zhi_d};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            const TuneChoice &c = it->second;
            if (h->zsolver <= 1) {
                h->zsolver = c.zsolver;
                h->ztiles = c.ztiles;
            }
            for (int i = 0; i < 4; i++) h->line_var[i] = c.var[i];
            h->tuned = true;
            return PP_SUCCESS;
        }
    }
    const size_t n = (size_t)h->g.nx * h->g.ny * h->g.nz;
    double *buf = nullptr;
    if (hipMalloc((void **)&buf, n * sizeof(double)) != hipSuccess) {
        // no room for a scratch field: keep the default configuration (tuning is an
        // optimisation, never a reason for the solve to fail)
        (void)hipGetLastError();
        h->tuned = true;
        return PP_SUCCESS;
    }
    int rc = launch_fill_uniform(buf, (long)n, 0x7475ull, 0, h->stream) == hipSuccess ? PP_SUCCESS : PP_HIP_ERROR;
    const int saved_force = h->force_var;
    h->force_var = -1;
    for (int i = 0; i < 4; i++) h->line_var[i] = 0;
    double st[5], tot, best = 1e30;
    int bz = h->zsolver, bt = h->ztiles;
    if (!rc && h->zsolver <= 1) {
        const int cand[3][2] = {{0, 1}, {0, 0}, {1, 0}};
        for (const auto &c : cand) {
            h->zsolver = c[0];
            h->ztiles = c[1];
            if ((c[1] && !use_ztiles(h)) || !fp64_z_supported(c[0], h->g.nz)) continue;
            if ((rc = tune_time(h, buf, 3, st, tot))) break;
            if (tot < best) {
                best = tot;
                bz = c[0];
                bt = c[1];
            }
        }
        h->zsolver = bz;
        h->ztiles = bt;
    }
    // line shapes: every combination of the stages' two shapes (at most 16), timed round-robin
    // over three rounds (best whole-solve time each, so drift spreads over all of them and
    // the effect of a stage's output order on the next stage counts); the default
    // combination is kept unless another is more than 0.5% faster
    const int len[4] = {h->g.nx, h->g.ny, h->g.ny, h->g.nx};
    int stages[4], nv = 0;
    for (int i = 0; i < 4; i++)
        if (line_has_variant(len[i])) stages[nv++] = i;
    const int ncomb = 1 << nv;
    std::vector<double> tc((size_t)ncomb, 1e30);
    for (int round = 0; round < 3 && !rc; round++) {
        for (int comb = 0; comb < ncomb && !rc; comb++) {
            for (int i = 0; i < 4; i++) h->line_var[i] = 0;
            for (int k = 0; k < nv; k++) h->line_var[stages[k]] = (comb >> k) & 1;
            rc = tune_time(h, buf, 1, st, tot);
            tc[comb] = std::min(tc[comb], tot);
        }
    }
    int bestc = 0;
    for (int comb = 1; comb < ncomb; comb++)
        if (tc[comb] < tc[bestc]) bestc = comb;
    if (tc[bestc] > 0.995 * tc[0]) bestc = 0;
    for (int i = 0; i < 4; i++) h->line_var[i] = 0;
    for (int k = 0; k < nv; k++) h->line_var[stages[k]] = (bestc >> k) & 1;
    h->force_var = saved_force;
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(buf);
    if (rc) return rc;
    h->tuned = true;
    TuneChoice c{h->zsolver, h->ztiles, {h->line_var[0], h->line_var[1], h->line_var[2], h->line_var[3]}};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = c;
    return PP_SUCCESS;
}

ZArgs z_args(pp_solver *h, void *data)
{
    ZArgs z{};
    z.data = data;
    z.SX = h->g.nx;
    z.SY = h->g.ny;
    z.nz = h->g.nz;
    z.lamx = h->d_lamx;
    z.lamy = h->d_lamy;
    z.Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
    z.Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
    z.zlo_dirichlet = h->g.zlo_d;
    z.zhi_dirichlet = h->g.zhi_d;
    z.zero_mode = h->zero_mode;
    z.tau = 0.0;
    z.pinned_zero = 0;
    return z;
}

// FP32 z tiles of the mixed form: the y transforms write and read 32-column tiles (128-byte
// rows, whole rows per block) of 16 J rows, so the FP32 z solve streams contiguous tiles
// instead of one 128-byte row from each of nz planes.  Compiled y lengths whose 32-column
// block fits (1024-point lines in their half-thread shape), columns up to 1024 rows (the FP64
// re-solve of near-singular columns works on 16-column tile halves up to 512 rows and column
// by column beyond).
bool use_ztiles_f32(const pp_solver *h)
{
    return h->ztiles && h->wtiles && !h->g.zper && line_tiled32_f32(h->g.ny) && zsolve_supported_f32(h->g.nz);
}

// y shape of the FP32 tiled stages: PP_OPT_LINE_VARIANT when forced, else the tuner's
int mixed_yvar(const pp_solver *h) { return h->force_var >= 0 ? h->force_var : h->wvar; }

int ztile_rows_f32(int nz) { return (nz + 15) / 16 * 16; }

// Buffers and the near-singular column list of the mixed form.  The list uses the same
// eigenvalue arithmetic as the kernels (no contraction) so both sides agree on it.
int ensure_mixed(pp_solver *h)
{
    if (h->mixed_ready) return PP_SUCCESS;
    if (use_ztiles_f32(h) && !h->d_wt) {
        const size_t elems = ((size_t)h->g.nx + 31) / 32 * 32 * h->g.ny * ztile_rows_f32(h->g.nz);
        if (hipMalloc((void **)&h->d_wt, elems * sizeof(float)) != hipSuccess) {
            h->d_wt = nullptr;
            return PP_ALLOCATION_FAIL;
        }
        // padding rows and columns are never written by the transforms; zero keeps them finite
        // through the z solve (identity rows)
        PP_HIP_CHECK(hipMemset(h->d_wt, 0, elems * sizeof(float)));
    }
    // size-dependent buffers: allocated once per handle (sizes are fixed at pp_create) and
    // freed only by pp_destroy, so no pointer handed to a solve can be released under it
    const size_t N = (size_t)h->g.nx * h->g.ny * h->g.nz;
    if (!h->d_w && hipMalloc((void **)&h->d_w, N * sizeof(float)) != hipSuccess) return PP_ALLOCATION_FAIL;
    // the solution's FP32 parts while it is refined (u0, then the sum of the corrections
    // before the last): the last store writes u = u0 + csum + c in FP64
    if (h->refine > 0 && !h->d_u0 && hipMalloc((void **)&h->d_u0, N * sizeof(float)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    if (h->refine > 1 && !h->d_cacc && hipMalloc((void **)&h->d_cacc, N * sizeof(float)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    if (!h->d_line_sums &&
        hipMalloc((void **)&h->d_line_sums, (size_t)h->g.ny * h->g.nz * sizeof(double)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    if (!h->d_plane_add &&
        hipMalloc((void **)&h->d_plane_add, (size_t)(h->g.nz + 2) * sizeof(double)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    if (!h->d_upart &&
        hipMalloc((void **)&h->d_upart, (size_t)residual_parts(h->g.nx, h->g.ny) * sizeof(double)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    // the near-singular column list depends on the tables, tau and the zero-mode rule
    if (h->d_fix) {
        PP_HIP_CHECK(hipStreamSynchronize(h->stream));
        PP_HIP_CHECK(hipFree(h->d_fix));
        h->d_fix = nullptr;
    }
    h->nfix = 0;
    const double Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
    const double Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
    const std::vector<double> &lx = h->lamx_h, &ly = h->lamy_h;
    std::vector<int4> fix;
    h->pinned00 = false;
    for (int sy = 0; sy < h->g.ny; sy++)
        for (int sx = 0; sx < h->g.nx; sx++) {
            const double eig = column_eig(lx, ly, sx, sy, Cx, Cy);
            bool pin = std::fabs(eig + 2.) < 1.e-10;
            if (pin && h->zero_mode == 1 && (h->g.zlo_d || h->g.zhi_d)) pin = false;
            if (pin) {
                h->pinned00 = true;
                continue;
            }
            // (periodic z computes every column in FP64: nothing to re-solve)
            if (!h->g.zper && std::fabs(eig + 2.) < h->tau) fix.push_back(make_int4(sx, sy, 0, 0));
        }
    h->nfix = (int)fix.size();
    if (h->nfix) {
        if (hipMalloc((void **)&h->d_fix, fix.size() * sizeof(int4)) != hipSuccess) return PP_ALLOCATION_FAIL;
        PP_HIP_CHECK(hipMemcpy(h->d_fix, fix.data(), fix.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    // the same columns as 16-column tiles for the FP64 tile re-solve
    if (h->d_fixtiles) {
        PP_HIP_CHECK(hipFree(h->d_fixtiles));
        h->d_fixtiles = nullptr;
    }
    std::vector<ZItem> tiles;
    for (const int4 &c : fix) {
        const ZItem it = {0, c.x / 16, c.y};
        if (tiles.empty() || tiles.back().bx != it.bx || tiles.back().by != it.by) {
            bool seen = false;
            for (const ZItem &o : tiles) seen = seen || (o.bx == it.bx && o.by == it.by);
            if (!seen) tiles.push_back(it);
        }
    }
    h->nfixtiles = (int)tiles.size();
    if (h->nfixtiles) {
        if (hipMalloc((void **)&h->d_fixtiles, tiles.size() * sizeof(ZItem)) != hipSuccess) return PP_ALLOCATION_FAIL;
        PP_HIP_CHECK(hipMemcpy(h->d_fixtiles, tiles.data(), tiles.size() * sizeof(ZItem), hipMemcpyHostToDevice));
    }
    {
        // rewritten whenever the tables change (it points at the eigenvalue factors)
        ZSolve zs{};
        zs.offset = 0;
        zs.SX = h->g.nx;
        zs.SY = h->g.ny;
        zs.nz = h->g.nz;
        zs.zlo_dirichlet = h->g.zlo_d;
        zs.zhi_dirichlet = h->g.zhi_d;
        zs.Cx = Cx;
        zs.Cy = Cy;
        zs.lamx = h->d_lamx;
        zs.lamy = h->d_lamy;
        if (!h->d_zself && hipMalloc((void **)&h->d_zself, sizeof(ZSolve)) != hipSuccess) return PP_ALLOCATION_FAIL;
        PP_HIP_CHECK(hipMemcpy(h->d_zself, &zs, sizeof(ZSolve), hipMemcpyHostToDevice));
    }
    h->mixed_ready = true;
    return PP_SUCCESS;
}

// one FP32 solve of the field in d_w, in place (forward x already applied by the caller
// when `x_forward_done`)
int mixed_inner(pp_solver *h)
{
    hipStream_t s = h->stream;
    int rc;
    const bool zt = use_ztiles_f32(h) && h->d_wt;
    float *zdata = zt ? h->d_wt : h->d_w;
    const int trows = ztile_rows_f32(h->g.nz);
    {
        LineArgs a = line_args(h, 1, nullptr);
        a.in = h->d_w;
        a.out = zdata;
        if (zt) {
            a.ztiled = ZT_OUT;
            a.tnz = trows;
            a.tile_w = 32;
            a.variant = mixed_yvar(h);
        }
        PP_HIP_CHECK(launch_line(FORM_F32, h->g.kind[1], h->g.ny, 1, 1, a, s));
    }
    if ((rc = mark(h))) return rc;
    if (h->g.zper) {
        // periodic z: FP64 arithmetic on the FP32 field, singular column zeroed
        if ((rc = run_zperiodic(h, h->d_w, true))) return rc;
        if ((rc = mark(h))) return rc;
    } else {
        ZArgs z = z_args(h, zdata);
        if (zt) {
            z.tiled = 1;
            z.tile_w = 32;
            z.tile_rows = trows;
        }
        z.tau = h->tau;
        z.pinned_zero = 1;
        PP_HIP_CHECK(launch_zsolve_pcr_f32(h->g.nz, z, s));
        if ((rc = mark(h))) return rc;
        if (zfix_tiles_supported(h->g.nz)) {
            ZArgs zf = z;
            zf.items = h->d_fixtiles;
            zf.solves = h->d_zself;
            zf.nitems = h->nfixtiles;
            PP_HIP_CHECK(launch_zfix_tiles_f32(h->g.nz, zf, s));
        } else {
            PP_HIP_CHECK(launch_zfix64_f32(z, h->d_fix, h->nfix, h->g.nz, s));
        }
    }
    if ((rc = mark(h))) return rc;
    {
        LineArgs a = line_args(h, 1, nullptr);
        a.in = zdata;
        a.out = h->d_w;
        if (zt) {
            a.ztiled = ZT_IN;
            a.tnz = trows;
            a.tile_w = 32;
            a.variant = mixed_yvar(h);
        }
        PP_HIP_CHECK(launch_line(FORM_F32, h->g.kind[1], h->g.ny, 1, 0, a, s));
    }
    return mark(h);
}

// Mixed precision: FP32 transforms and z solves (FP64 twiddles, eigenvalues and
// near-singular columns), the pinned (0,0) column exactly in FP64, then `refine` steps of
// u += solve32(f_eff - L u) with the residual in FP64.  The solution is held as FP32 parts
// while it is refined -- u = u0 + csum + c_k, u0 the initial FP32 solve, csum the corrections
// before the last (each summed into it in FP32), the residual reading the parts and adding them
// in FP64 -- and only the last store writes u in FP64 (u0 + csum + c_R); csum's FP32 rounding
// (eps32 |c| ~ 1e-14 |u| after the first correction) is what the next step corrects.  f and u
// must not alias.
// the correction step runs fused (pp_resx.h) where the x lines have a compiled plan and ny
// is a multiple of the fused kernel's lines per block (PP_OPT_FUSE_RESIDUAL, default on)
static bool mixed_fused(const pp_solver *h)
{
    const int lines = resx_lines(h->g.nx);
    return h->fuse_res && line_compiled(h->g.nx) && lines > 0 && h->g.ny % lines == 0;
}

int run_mixed(pp_solver *h, const double *f, double *u)
{
    hipStream_t s = h->stream;
    int rc = ensure_mixed(h);
    if (rc) return rc;
    if ((rc = begin_timing(h))) return rc;
    const int R = h->refine;
    // initial solve: x forward reads the FP64 right-hand side
    {
        LineArgs a = line_args(h, 0, nullptr);
        a.in = f;
        a.out = h->d_w;
        a.in_scale = h->g.dxyz[2];
        a.dmask = h->g.dmask;
        a.line_sums = h->pinned00 ? h->d_line_sums : nullptr;
        PP_HIP_CHECK(launch_line(FORM_F64_TO_F32, h->g.kind[0], h->g.nx, 0, 1, a, s));
    }
    if ((rc = mark(h))) return rc;
    if ((rc = mixed_inner(h))) return rc;
    if (h->pinned00) {
        const double b0 = h->g.zlo_d ? -2.0 - 1. : -2.0 + 1.;
        const double bn = h->g.zhi_d ? -2.0 - 1. : -2.0 + 1.;
        PP_HIP_CHECK(launch_mode00(h->d_line_sums, h->g.ny, h->g.nz, b0, bn, 1. / h->g.scale, h->d_plane_add, nullptr, 0, 0,
                                   h->g.zper, s));
    }
    {
        // u0 in FP32 while refinement follows, else the FP64 solution
        LineArgs a = line_args(h, 0, nullptr);
        a.in = h->d_w;
        a.out = R > 0 ? (void *)h->d_u0 : (void *)u;
        a.out_scale = 1. / (h->g.scale);
        a.plane_add = h->pinned00 ? h->d_plane_add : nullptr;
        a.accumulate = 0;
        PP_HIP_CHECK(launch_line(R > 0 ? FORM_F32 : FORM_F32_TO_F64, h->g.kind[0], h->g.nx, 0, 0, a, s));
    }
    if ((rc = mark(h))) return rc;
    const double Cx = (h->g.dxyz[2] / h->g.dxyz[0]) * (h->g.dxyz[2] / h->g.dxyz[0]);
    const double Cy = (h->g.dxyz[2] / h->g.dxyz[1]) * (h->g.dxyz[2] / h->g.dxyz[1]);
    const bool fused = mixed_fused(h);
    for (int it = 0; it < R; it++) {
        const bool last = it == R - 1;
        ResArgs r{};
        r.f = f;
        r.u = nullptr;
        r.u0 = h->d_u0;
        r.u1 = h->d_cacc;
        r.r = h->d_w;
        r.nx = h->g.nx;
        r.ny = h->g.ny;
        r.nz = h->g.nz;
        r.c3 = h->g.dxyz[2];
        r.Cx = Cx;
        r.Cy = Cy;
        for (int i = 0; i < 6; i++) r.bc[i] = h->g.bc[i];
        r.dtab = h->d_dtab;
        r.dmask = h->g.dmask;
        r.pin_defect = h->pinned00 ? h->d_plane_add + h->g.nz : nullptr;
        r.upart = h->pinned00 ? h->d_upart : nullptr;
        const int src = it == 0 ? RES_U0 : RES_U0_U1;
        LineArgs a = line_args(h, 0, nullptr);
        a.in = h->d_w;
        a.out = h->d_w;
        a.line_sums = h->pinned00 ? h->d_line_sums : nullptr;
        if (fused) {
            // residual computed inside the x forward: one kernel, r never stored
            PP_HIP_CHECK(launch_resx(h->g.kind[0], h->g.nx, src, a, r, nullptr, s));
        } else {
            PP_HIP_CHECK(launch_residual_f32(r, src, s));
            if ((rc = mark(h))) return rc;
            PP_HIP_CHECK(launch_line(FORM_F32, h->g.kind[0], h->g.nx, 0, 1, a, s));
        }
        if ((rc = mark(h))) return rc;
        if ((rc = mixed_inner(h))) return rc;
        if (h->pinned00) {
            // the FP32 passes leak rounding into the plane means of u, which the zeroed
            // (0,0) column of the FP32 correction cannot see: correct that column in FP64
            const double b0 = h->g.zlo_d ? -2.0 - 1. : -2.0 + 1.;
            const double bn = h->g.zhi_d ? -2.0 - 1. : -2.0 + 1.;
            PP_HIP_CHECK(launch_mode00(h->d_line_sums, h->g.ny, h->g.nz, b0, bn, 1. / h->g.scale, h->d_plane_add, h->d_upart,
                                       fused ? h->g.ny / resx_lines(h->g.nx) : residual_parts(h->g.nx, h->g.ny), 1,
                                       h->g.zper, s));
        }
        {
            LineArgs a = line_args(h, 0, nullptr);
            a.in = h->d_w;
            a.out_scale = 1. / (h->g.scale);
            a.plane_add = h->pinned00 ? h->d_plane_add : nullptr;
            if (last) {
                // u = u0 + csum + c, in FP64
                a.out = u;
                a.accumulate = 0;
                a.acc0 = h->d_u0;
                a.acc1 = it > 0 ? h->d_cacc : nullptr;
                PP_HIP_CHECK(launch_line(FORM_F32_TO_F64, h->g.kind[0], h->g.nx, 0, 0, a, s));
            } else {
                // csum = c1, then csum += c
                a.out = h->d_cacc;
                a.accumulate = it > 0 ? 1 : 0;
                PP_HIP_CHECK(launch_line(FORM_F32, h->g.kind[0], h->g.nx, 0, 0, a, s));
            }
        }
        if ((rc = mark(h))) return rc;
    }
    end_timing(h);
    return PP_SUCCESS;
}

// Autotuner of the mixed form: FP32 z tiles between the y transforms and the z solve or the
// work field in place, and with tiles the FP32 y shape (32 threads or half of them per
// column).  Whole mixed solves on scratch fields, best of three, timed round-robin; the
// default (tiles, default shape) is kept unless another is more than 0.5% faster.  Cached per
// process and shape.
std::map<std::array<int, 7>, std::array<int, 2>> g_mtune;

int autotune_mixed(pp_solver *h)
{
    const std::array<int, 7> key = {h->g.nx, h->g.ny, h->g.nz, h->g.kind[0], h->g.kind[1], h->g.zlo_d, h->g.zhi_d};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_mtune.find(key);
        if (it != g_mtune.end()) {
            h->wtiles = it->second[0];
            h->wvar = it->second[1];
            h->mtuned = true;
            return PP_SUCCESS;
        }
    }
    h->wtiles = 1;
    h->wvar = 0;
    std::vector<std::array<int, 2>> cand = {{1, 0}};
    if (use_ztiles_f32(h)) {
        if (line_tiled32_f32_variant(h->g.ny)) cand.push_back({1, 1});
        cand.push_back({0, 0});
    }
    const size_t n = (size_t)h->g.nx * h->g.ny * h->g.nz;
    double *f = nullptr, *u = nullptr;
    if (cand.size() < 2 || hipMalloc((void **)&f, n * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&u, n * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        if (f) (void)hipFree(f);
        h->mtuned = true;  // nothing to choose, or no room for scratch fields: the default
        return PP_SUCCESS;
    }
    hipEvent_t e0, e1;
    int rc = PP_SUCCESS;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PP_HIP_ERROR;
    if (!rc && launch_fill_uniform(f, (long)n, 0x6d74ull, 0, h->stream) != hipSuccess) rc = PP_HIP_ERROR;
    std::vector<double> best(cand.size(), 1e30);
    for (int round = 0; round < 4 && !rc; round++)
        for (size_t c = 0; c < cand.size() && !rc; c++) {
            h->wtiles = cand[c][0];
            h->wvar = cand[c][1];
            h->mixed_ready = false;  // the tile buffer follows the choice (set up untimed)
            if ((rc = ensure_mixed(h))) break;
            if (round == 0 && (rc = run_mixed(h, f, u))) break;  // warm-up
            if (hipEventRecord(e0, h->stream) != hipSuccess) rc = PP_HIP_ERROR;
            if (!rc) rc = run_mixed(h, f, u);
            if (!rc && (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
                rc = PP_HIP_ERROR;
            float t = 0.f;
            if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = PP_HIP_ERROR;
            best[c] = std::min(best[c], (double)t);
        }
    size_t bc = 0;
    for (size_t c = 1; c < cand.size(); c++)
        if (best[c] < best[bc]) bc = c;
    if (best[bc] > 0.995 * best[0]) bc = 0;
    h->wtiles = cand[bc][0];
    h->wvar = cand[bc][1];
    h->mixed_ready = false;
    (void)hipStreamSynchronize(h->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(f);
    (void)hipFree(u);
    if (rc) return rc;
    h->mtuned = true;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_mtune[key] = cand[bc];
    return PP_SUCCESS;
}

// dispatch on the precision form; rhs and out may alias (the mixed form then works on a copy)
int run_solve(pp_solver *h, const double *rhs, double *out);

// a solve through the graph path when enabled: the first solve on a pointer pair runs
// eagerly (tables, buffers, the tuner), the second is captured -- kernel launches only, no
// memset or copy is recorded -- and later ones replay it
int solve_graphed(pp_solver *h, const double *rhs, double *out)
{
    const bool graphable = h->graph_mode && !h->timing && h->stream != nullptr &&
                           !(h->precision == 0 && h->autotune && !h->tuned) &&
                           !(h->precision == 1 && h->autotune && !h->mtuned) && !(h->precision == 1 && rhs == out);
    if (!graphable) {
        if (h->graph) drop_graph(h);
        return run_solve(h, rhs, out);
    }
    if (h->g_rhs != rhs || h->g_out != out) {
        drop_graph(h);
        h->g_rhs = rhs;
        h->g_out = out;
    }
    if (h->graph) {
        PP_HIP_CHECK(hipGraphLaunch(h->graph, h->stream));
        return PP_SUCCESS;
    }
    if (h->g_eager == 0) {
        const int rc = run_solve(h, rhs, out);
        if (!rc) h->g_eager = 1;
        return rc;
    }
    hipGraph_t g = nullptr;
    PP_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
    const int rc = run_solve(h, rhs, out);
    const hipError_t ec = hipStreamEndCapture(h->stream, &g);
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : PP_HIP_ERROR;
    }
    const hipError_t ei = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        h->graph = nullptr;
        return PP_HIP_ERROR;
    }
    PP_HIP_CHECK(hipGraphLaunch(h->graph, h->stream));
    return PP_SUCCESS;
}

int run_solve(pp_solver *h, const double *rhs, double *out)
{
    if (h->precision == 0) {
        if (h->autotune && !h->tuned && !h->g.zper) {
            const int rc = autotune(h);
            if (rc) return rc;
        }
        return run_fp64(h, rhs, out);
    }
    if (h->autotune && !h->mtuned && !h->g.zper) {
        const int rc = autotune_mixed(h);
        if (rc) return rc;
    }
    if (rhs == out) {
        const size_t bytes = (size_t)h->g.nx * h->g.ny * h->g.nz * sizeof(double);
        if (!h->d_aux && hipMalloc((void **)&h->d_aux, bytes) != hipSuccess) return PP_ALLOCATION_FAIL;
        PP_HIP_CHECK(hipMemcpyAsync(h->d_aux, rhs, bytes, hipMemcpyDeviceToDevice, h->stream));
        return run_mixed(h, h->d_aux, out);
    }
    return run_mixed(h, rhs, out);
}

}  // namespace

namespace pp {
int run_fp64(pp_solver *h, const double *rhs, double *field) { return run_fp64_stages(h, rhs, field); }
}  // namespace pp

extern "C" {

int pp_create(int nx, int ny, int nz, pp_handle *out)
{
    if (!out) return PP_INPUT_ARGS;
    *out = nullptr;
    if (nx < 2 || ny < 2 || nz < 4) return PP_INPUT_ARGS;
    if (!size_supported(nx) || !size_supported(ny)) return PP_UNSUPPORTED_SIZE;
    if (!zsolve_supported_f64(nz) && !zsolve_supported_f32(nz) && !zperiodic_supported(nz)) return PP_UNSUPPORTED_SIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PP_NO_DEVICE;
    pp_solver *h = new (std::nothrow) pp_solver();
    if (!h) return PP_ALLOCATION_FAIL;
    h->g.nx = nx;
    h->g.ny = ny;
    h->g.nz = nz;
    const size_t bytes = (size_t)nx * ny * nz * sizeof(double);
    if (hipMalloc((void **)&h->d_field, bytes) != hipSuccess) {
        delete h;
        return PP_ALLOCATION_FAIL;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipFree(h->d_field);
        delete h;
        return PP_HIP_ERROR;
    }
    h->own_stream = true;
    *out = h;
    return PP_SUCCESS;
}

int pp_destroy(pp_handle h)
{
    if (!h) return PP_INVALID_HANDLE;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graph(h);
    free_tables(h);
    free_mixed(h);
    if (h->d_zt) (void)hipFree(h->d_zt);
    if (h->d_partials) (void)hipFree(h->d_partials);
    if (h->d_field) (void)hipFree(h->d_field);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    h->ev.clear();
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return PP_SUCCESS;
}

int pp_assign_boundary(pp_handle h, const char *bc)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    if (!bc) return PP_INPUT_ARGS;
    const int rc = setup_boundary(h->g, bc);
    if (rc) return rc;
    h->bc_set = true;
    h->tables_ready = false;
    return PP_SUCCESS;
}

int pp_set_box(pp_handle h, const double *X)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    if (!X) return PP_INPUT_ARGS;
    // an invalid box leaves the handle's box (and its tables) as they were
    if (!(X[1] > X[0] && X[3] > X[2] && X[5] > X[4])) return PP_INPUT_ARGS;
    for (int i = 0; i < 6; i++) h->g.X[i] = X[i];
    h->box_set = true;
    h->tables_ready = false;
    if (h->coords_set) return pp_set_coords(h, 2);
    return PP_SUCCESS;
}

int pp_set_coords(pp_handle h, int dir)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    if (dir < 0 || dir > 2) return PP_INPUT_ARGS;
    setup_coords(h->g);
    h->coords_set = true;
    h->tables_ready = false;
    return PP_SUCCESS;
}

int pp_assign_rhs(pp_handle h, const double *rhs)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!rhs) return PP_INPUT_ARGS;
    const size_t bytes = (size_t)h->g.nx * h->g.ny * h->g.nz * sizeof(double);
    PP_HIP_CHECK(hipMemcpyAsync(h->d_field, rhs, bytes, hipMemcpyHostToDevice, h->stream));
    PP_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->rhs_set = true;
    h->trig_part = -1;
    return PP_SUCCESS;
}

int pp_assign_rhs_device(pp_handle h, const double *d_rhs)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!d_rhs) return PP_INPUT_ARGS;
    const size_t bytes = (size_t)h->g.nx * h->g.ny * h->g.nz * sizeof(double);
    PP_HIP_CHECK(hipMemcpyAsync(h->d_field, d_rhs, bytes, hipMemcpyDeviceToDevice, h->stream));
    h->rhs_set = true;
    h->trig_part = -1;
    return PP_SUCCESS;
}

int pp_pittpack(pp_handle h)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!h->rhs_set) return PP_CALL_ORDER;
    int rc = ensure_ready(h);
    if (rc) return rc;
    rc = run_solve(h, h->d_field, h->d_field);
    if (rc) return rc;
    PP_HIP_CHECK(hipStreamSynchronize(h->stream));
    h->rhs_set = false;  // the field now holds the solution
    return PP_SUCCESS;
}

int pp_solve_device(pp_handle h, double *d_field)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!d_field) return PP_INPUT_ARGS;
    int rc = ensure_ready(h);
    if (rc) return rc;
    return solve_graphed(h, d_field, d_field);
}

int pp_solve_device_oop(pp_handle h, const double *d_rhs, double *d_u)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!d_rhs || !d_u) return PP_INPUT_ARGS;
    int rc = ensure_ready(h);
    if (rc) return rc;
    return solve_graphed(h, d_rhs, d_u);
}

int pp_get_solution(pp_handle h, double *u)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!u) return PP_INPUT_ARGS;
    const size_t bytes = (size_t)h->g.nx * h->g.ny * h->g.nz * sizeof(double);
    PP_HIP_CHECK(hipMemcpyAsync(u, h->d_field, bytes, hipMemcpyDeviceToHost, h->stream));
    PP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return PP_SUCCESS;
}

int pp_get_solution_device(pp_handle h, double *d_u)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!d_u) return PP_INPUT_ARGS;
    const size_t bytes = (size_t)h->g.nx * h->g.ny * h->g.nz * sizeof(double);
    PP_HIP_CHECK(hipMemcpyAsync(d_u, h->d_field, bytes, hipMemcpyDeviceToDevice, h->stream));
    return PP_SUCCESS;
}

int pp_set_stream(pp_handle h, void *stream)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    PP_HIP_CHECK(hipStreamSynchronize(h->stream));
    if (h->own_stream) {
        (void)hipStreamDestroy(h->stream);
        h->own_stream = false;
    }
    h->stream = reinterpret_cast<hipStream_t>(stream);
    return PP_SUCCESS;
}

int pp_set_option(pp_handle h, int key, long long value)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    switch (key) {
    case PP_OPT_ZERO_MODE:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->zero_mode = (int)value;
        h->mixed_ready = false;
        h->zthomas_ready = false;
        return PP_SUCCESS;
    case PP_OPT_ZSOLVER:
        if (value < 0 || value > 4) return PP_INPUT_ARGS;
        h->zsolver = (int)value;
        return PP_SUCCESS;
    case PP_OPT_ZLAYOUT:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->ztiles = (int)value;
        h->mixed_ready = false;
        return PP_SUCCESS;
    case PP_OPT_AUTOTUNE:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->autotune = (int)value;
        h->tuned = false;
        h->mtuned = false;
        if (!value) {
            for (int i = 0; i < 4; i++) h->line_var[i] = 0;
            h->wtiles = 1;
            h->wvar = 0;
            h->mixed_ready = false;
        }
        return PP_SUCCESS;
    case PP_OPT_LINE_VARIANT:
        if (value < -1 || value > 1) return PP_INPUT_ARGS;
        h->force_var = (int)value;
        return PP_SUCCESS;
    case PP_OPT_GRAPH:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->graph_mode = (int)value;
        return PP_SUCCESS;
    case PP_OPT_PRECISION:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->precision = (int)value;
        return PP_SUCCESS;
    case PP_OPT_IO_FORMAT:
        if (value != PP_IO_HDF5 && value != PP_IO_RAW) return PP_INPUT_ARGS;
        h->io_format = (int)value;
        return PP_SUCCESS;
    case PP_OPT_FUSE_RESIDUAL:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        h->fuse_res = (int)value;
        return PP_SUCCESS;
    case PP_OPT_REFINE:
        if (value < 0 || value > 16) return PP_INPUT_ARGS;
        h->refine = (int)value;
        h->mixed_ready = false;  // the FP32 solution parts it needs
        return PP_SUCCESS;
    case PP_OPT_TIMING: {
        if (value < 0 || value > 1024) return PP_INPUT_ARGS;
        PP_HIP_CHECK(hipStreamSynchronize(h->stream));
        for (auto &e : h->ev)
            if (e) (void)hipEventDestroy(e);
        h->ev.assign((size_t)pp_solver::kMaxMarks * value, nullptr);
        for (auto &e : h->ev)
            if (hipEventCreate(&e) != hipSuccess) return PP_HIP_ERROR;
        h->ev_slots = (int)value;
        h->ev_next = 0;
        h->ev_used = 0;
        h->timing = value != 0;
        return PP_SUCCESS;
    }
    default: return PP_INPUT_ARGS;
    }
}

int pp_get_tuning(pp_handle h, int *out, int n)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!out || n < 7) return PP_INPUT_ARGS;
    out[0] = h->tuned ? 1 : 0;
    out[1] = h->zsolver;
    out[2] = use_ztiles(h) ? 1 : 0;
    for (int i = 0; i < 4; i++) out[3 + i] = stage_var(h, i, i == 0 || i == 3 ? h->g.nx : h->g.ny);
    if (n >= 9) {
        // mixed form: FP32 z tiles in use, and the tiled y shape
        out[7] = use_ztiles_f32(h) ? 1 : 0;
        out[8] = out[7] && line_tiled32_f32_variant(h->g.ny) ? mixed_yvar(h) : 0;
    }
    return PP_SUCCESS;
}

int pp_get_kernel_times(pp_handle h, float *ms, int max, int *count)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!ms || !count || max < 0) return PP_INPUT_ARGS;
    if (!h->timing || h->ev_used == 0 || h->marks_last < 2) return PP_CALL_ORDER;
    PP_HIP_CHECK(hipStreamSynchronize(h->stream));
    const int n = h->marks_last - 1;
    std::vector<double> acc(n, 0.0);
    for (int slot = 0; slot < h->ev_used; slot++) {
        hipEvent_t *ev = &h->ev[(size_t)pp_solver::kMaxMarks * slot];
        for (int i = 0; i < n; i++) {
            float t = 0.f;
            PP_HIP_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            acc[i] += t;
        }
    }
    *count = n;
    for (int i = 0; i < n && i < max; i++) ms[i] = (float)(acc[i] / h->ev_used);
    h->ev_used = 0;
    h->ev_next = 0;
    return PP_SUCCESS;
}

int pp_get_stage_times(pp_handle h, float *ms5)
{
    if (!ms5) return PP_INPUT_ARGS;
    float ms[pp_solver::kMaxMarks];
    int n = 0;
    int rc = pp_get_kernel_times(h, ms, pp_solver::kMaxMarks, &n);
    if (rc) return rc;
    if (n < 5) return PP_CALL_ORDER;
    for (int i = 0; i < 5; i++) ms5[i] = ms[i];
    return PP_SUCCESS;
}

int pp_set_option_double(pp_handle h, int key, double value)
{
    if (!h) return PP_INVALID_HANDLE;
    drop_graph(h);
    if (key == PP_OPT_MG_TOL) {
        if (!(value > 0.0 && value < 1.0)) return PP_INPUT_ARGS;
        h->mg_tol = value;
        return PP_SUCCESS;
    }
    if (key == PP_OPT_TAU) {
        if (!(value >= 0.0 && value < 4.0)) return PP_INPUT_ARGS;
        h->tau = value;
        h->mixed_ready = false;
        return PP_SUCCESS;
    }
    if (key == PP_OPT_ZTHOMAS_TAU) {
        if (!(value >= 0.0 && value < 4.0)) return PP_INPUT_ARGS;
        h->zthomas_tau = value;
        h->zthomas_ready = false;
        return PP_SUCCESS;
    }
    return PP_INPUT_ARGS;
}

int pp_get_device_field(pp_handle h, double **d_field)
{
    if (!h) return PP_INVALID_HANDLE;
    if (!d_field) return PP_INPUT_ARGS;
    *d_field = h->d_field;
    return PP_SUCCESS;
}

const char *pp_get_error_enum(int code)
{
    switch (code) {
    case 0: return "SUCCESS";
    case 1: return "GPU_INIT_FAILURE";
    case 2: return "SOLVER_ARRAY_INCONSISTENT";
    case 3: return "MPI_INIT_FAIL";
    case 4: return "MPI_DUP_FAIL";
    case 5: return "COMSIZE_FAIL";
    case 6: return "MESH_DIVISIBLE";
    case 7: return "MPI_GET_RANK_FAIL";
    case 8: return "MPI_COMSIZE_FAIL";
    case 9: return "MPI_DIST_GTAPH_FAIL_X";
    case 10: return "MPI_DIST_GTAPH_FAIL_Y";
    case 11: return "GRAPH_CREATE_FAIL";
    case 12: return "MPI_INEIGHBOR_FAIL_ZX";
    case 13: return "MPI_INEIGHBOR_FAIL_XY";
    case 14: return "BLOCK_NUMBER_FAIL";
    case 15: return "ALLOCATION_FAIL";
    case 16: return "MPI_FINALIZE_FAIL";
    case 17: return "CONNECTIVITY_CONSTRUCTION_FAIL";
    case 18: return "THOMAS_FAIL";
    case 19: return "INPUT_ARGS";
    case 20: return "CUFFT_FAIL_X";
    case 21: return "CUFFT_FAIL_INV_X";
    case 22: return "CUFFT_FAIL_Y";
    case 23: return "CUFFT_FAIL_INV_Y";
    case PP_HIP_ERROR: return "HIP_ERROR";
    case PP_UNSUPPORTED_BC: return "UNSUPPORTED_BC";
    case PP_UNSUPPORTED_SIZE: return "UNSUPPORTED_SIZE";
    case PP_INVALID_HANDLE: return "INVALID_HANDLE";
    case PP_CALL_ORDER: return "CALL_ORDER";
    case PP_NO_DEVICE: return "NO_DEVICE";
    case PP_IO_FAIL: return "IO_FAIL";
    case PP_NO_HDF5: return "NO_HDF5";
    default: return "UNKNOWN_ERROR";
    }
}

int pp_size_supported(int n) { return size_supported(n) ? 1 : 0; }

int pp_selftest(void)
{
    int *d = nullptr;
    int host[6 * 64];
    if (hipMalloc((void **)&d, sizeof(host)) != hipSuccess) return PP_ALLOCATION_FAIL;
    hipLaunchKernelGGL(k_selftest_rot, dim3(1), dim3(64), 0, 0, d);
    if (hipGetLastError() != hipSuccess || hipMemcpy(host, d, sizeof(host), hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipFree(d);
        return PP_HIP_ERROR;
    }
    (void)hipFree(d);
    const int S[6] = {1, 2, 4, 8, -1, -8};
    int bad = 0;
    for (int r = 0; r < 6; r++)
        for (int lane = 0; lane < 64; lane++) {
            int row = lane & ~15, t = lane & 15;
            int expect = row + ((t - S[r]) & 15);
            if (host[r * 64 + lane] != expect) bad++;
        }
    return bad ? 1000 + bad : 0;
}

}  // extern "C"
