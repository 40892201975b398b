// pp_batch.hip -- batched form: many independent solves of uneven sizes and boundary
// conditions, each stage run as one launch per shape class over all members of the class.
//
// The reference solves one grid per PencilDcmp object (src/pencilDcmp.cpp:5801-6430); here
// the members' fields sit back to back in one device arena and every kernel launch covers
// the blocks of all members that share its template (line length and transform kind for
// the x and y stages, rows per lane for the z stage).  A block finds its member through a
// per-launch item table; per-member scalars (sizes, dz, 1/(NxNy), Dirichlet tables,
// eigenvalue factors) come from a per-member descriptor.  Kernels and arithmetic are those
// of the single-solve path, so a member's result does not depend on the batch around it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
#include <vector>

#include "pittpack_hip.h"
#include "pp_launch.h"
#include "pp_mixed.h"
#include "pp_setup.h"

using namespace pp;

#define PPB_CHECK(x)                                  \
    do {                                              \
        if ((x) != hipSuccess) return PP_HIP_ERROR;   \
    } while (0)

namespace {

struct BatchMember {
    SolveSetup g;
    long offset = 0;  // element offset of the member's field in the rhs / solution / work arenas
    long lam_off[2] = {0, 0};
    long dtab_off = 0;
    // mixed form
    bool pinned = false;  // pinned (0,0) column solved in FP64 by the plane-sum path
    long sums_off = 0, plane_off = 0, upart_off = 0;
};

// one launch of one stage: a shape class and its block items
struct ClassLaunch {
    int key = 0;   // x/y: line length N; z: rows per lane J
    int kind = 0;  // x/y: transform kind
    int nitems = 0;
    void *d_items = nullptr;
};

struct AxisDev {
    double *tw = nullptr, *A = nullptr, *B = nullptr, *C = nullptr, *D = nullptr;
};

template <typename T>
int upload_vec(T **dst, const std::vector<T> &v)
{
    *dst = nullptr;
    if (v.empty()) return PP_SUCCESS;
    if (hipMalloc((void **)dst, v.size() * sizeof(T)) != hipSuccess) return PP_ALLOCATION_FAIL;
    if (hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return PP_HIP_ERROR;
    return PP_SUCCESS;
}

}  // namespace

struct pp_batch_s {
    std::vector<BatchMember> mem;
    bool finalized = false;
    int zero_mode = 1;
    int precision = 0;  // 0 FP64, 1 FP32 + FP64 refinement (as PP_OPT_PRECISION)
    int refine = 2;
    double tau = 0.05;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // device state built by pp_batch_finalize
    double *d_rhs = nullptr, *d_sol = nullptr;  // right-hand sides (kept) and solutions
    float *d_w = nullptr;                       // FP32 work arena (mixed form)
    long arena_elems = 0;
    double *d_lam = nullptr, *d_dtab = nullptr;
    LineSolve *d_lsolve = nullptr;
    ZSolve *d_zsolve = nullptr;
    std::map<int, AxisDev> tables;  // per line length
    std::vector<ClassLaunch> lx, ly, lz;  // FP64 launches (x, y: items sized for FP64 blocks)
    std::vector<ClassLaunch> lx32, ly32;  // FP32 launches of the mixed form
    // mixed form
    double *d_sums = nullptr, *d_plane = nullptr, *d_upart = nullptr;
    int4 *d_fix = nullptr;
    int nfix = 0;
    Mode00Item *d_m00 = nullptr;
    int nm00 = 0;
    ResArgs *d_res = nullptr;
    int4 *d_res_items = nullptr;
    int nres_items = 0;
    int nz_max = 0;
    // timing: events between the stages of the last solves
    bool timing = false;
    static constexpr int kMarks = 64;
    std::vector<hipEvent_t> ev;
    int ev_slots = 0, ev_next = 0, ev_used = 0, marks_last = 0;
    hipEvent_t *ev_cur = nullptr;
    int marks_cur = 0;
};

namespace {

void free_device(pp_batch_s *b)
{
    auto F = [](void *p) {
        if (p) (void)hipFree(p);
    };
    for (void *p : {(void *)b->d_rhs, (void *)b->d_sol, (void *)b->d_w, (void *)b->d_lam, (void *)b->d_dtab,
                    (void *)b->d_lsolve, (void *)b->d_zsolve, (void *)b->d_sums, (void *)b->d_plane,
                    (void *)b->d_upart, (void *)b->d_fix, (void *)b->d_m00, (void *)b->d_res,
                    (void *)b->d_res_items})
        F(p);
    for (auto &kv : b->tables) {
        F(kv.second.tw);
        F(kv.second.A);
        F(kv.second.B);
        F(kv.second.C);
        F(kv.second.D);
    }
    b->tables.clear();
    for (auto *v : {&b->lx, &b->ly, &b->lz, &b->lx32, &b->ly32}) {
        for (auto &c : *v) F(c.d_items);
        v->clear();
    }
    b->d_rhs = b->d_sol = nullptr;
    b->d_w = nullptr;
    b->d_lam = b->d_dtab = nullptr;
    b->d_lsolve = nullptr;
    b->d_zsolve = nullptr;
    b->d_sums = b->d_plane = b->d_upart = nullptr;
    b->d_fix = nullptr;
    b->d_m00 = nullptr;
    b->d_res = nullptr;
    b->d_res_items = nullptr;
    b->nfix = b->nm00 = b->nres_items = 0;
    b->finalized = false;
}

int mark(pp_batch_s *b)
{
    if (b->ev_cur && b->marks_cur < pp_batch_s::kMarks) PPB_CHECK(hipEventRecord(b->ev_cur[b->marks_cur++], b->stream));
    return PP_SUCCESS;
}

int table_for(pp_batch_s *b, int N)
{
    if (b->tables.count(N)) return PP_SUCCESS;
    AxisTables t;
    make_axis_tables(N, t);
    AxisDev d;
    int rc;
    if ((rc = upload_vec(&d.tw, t.tw)) || (rc = upload_vec(&d.A, t.A)) || (rc = upload_vec(&d.B, t.B)) ||
        (rc = upload_vec(&d.C, t.C)) || (rc = upload_vec(&d.D, t.D)))
        return rc;
    b->tables[N] = d;
    return PP_SUCCESS;
}

// x (axis 0) or y (axis 1) stage: one class per (N, kind); block sizes of precision R
template <typename R>
int build_line_classes(pp_batch_s *b, int axis, std::vector<ClassLaunch> &out)
{
    std::map<std::pair<int, int>, std::vector<LineItem>> cls;
    for (int m = 0; m < (int)b->mem.size(); m++) {
        const SolveSetup &g = b->mem[m].g;
        const int N = axis == 0 ? g.nx : g.ny;
        const int kind = g.kind[axis];
        const int lines = line_block_lines<R>(N, axis);
        if (lines <= 0) return PP_UNSUPPORTED_SIZE;
        auto &items = cls[{N, kind}];
        if (axis == 0) {
            const long nl = (long)g.ny * g.nz;
            for (long bx = 0; bx < (nl + lines - 1) / lines; bx++) items.push_back({m, (int)bx, 0});
        } else {
            const int nbx = (g.nx + lines - 1) / lines;
            for (int z = 0; z < g.nz; z++)
                for (int bx = 0; bx < nbx; bx++) items.push_back({m, bx, z});
        }
    }
    for (auto &kv : cls) {
        ClassLaunch c;
        c.key = kv.first.first;
        c.kind = kv.first.second;
        c.nitems = (int)kv.second.size();
        int rc = upload_vec((LineItem **)&c.d_items, kv.second);
        if (rc) return rc;
        out.push_back(c);
        if ((rc = table_for(b, c.key))) return rc;
    }
    return PP_SUCCESS;
}

int build_z_classes(pp_batch_s *b)
{
    std::map<int, std::vector<ZItem>> cls;
    for (int m = 0; m < (int)b->mem.size(); m++) {
        const SolveSetup &g = b->mem[m].g;
        const int J = (g.nz + 15) / 16;
        auto &items = cls[J];
        const int nbx = (g.nx + 15) / 16;
        for (int sy = 0; sy < g.ny; sy++)
            for (int bx = 0; bx < nbx; bx++) items.push_back({m, bx, sy});
    }
    for (auto &kv : cls) {
        ClassLaunch c;
        c.key = kv.first;
        c.nitems = (int)kv.second.size();
        int rc = upload_vec((ZItem **)&c.d_items, kv.second);
        if (rc) return rc;
        b->lz.push_back(c);
    }
    return PP_SUCCESS;
}

// one stage of line transforms over every class: in/out arenas, precision form, flags
int run_line_stage(pp_batch_s *b, const std::vector<ClassLaunch> &cls, int form, int axis, int fwd, const void *in,
                   void *out, int bflags, double *sums, const double *plane, int accumulate)
{
    for (const ClassLaunch &c : cls) {
        const AxisDev &t = b->tables.at(c.key);
        LineArgs a{};
        a.in = in;
        a.out = out;
        a.tw = t.tw;
        a.A = t.A;
        a.B = t.B;
        a.C = t.C;
        a.D = t.D;
        a.in_scale = 1.0;
        a.out_scale = 1.0;
        a.items = (const LineItem *)c.d_items;
        a.solves = b->d_lsolve;
        a.nlines = c.nitems;
        a.bflags = bflags;
        a.line_sums = sums;
        a.plane_add = plane;
        a.accumulate = accumulate;
        PPB_CHECK(launch_line(form, c.kind, c.key, axis, fwd, a, b->stream));
    }
    return PP_SUCCESS;
}

// z stage over every rows-per-lane class; FP32 with the mixed form's column rules
int run_z_stage(pp_batch_s *b, bool f32)
{
    for (const ClassLaunch &c : b->lz) {
        ZArgs z{};
        z.data = f32 ? (void *)b->d_w : (void *)b->d_sol;
        z.zero_mode = b->zero_mode;
        z.items = (const ZItem *)c.d_items;
        z.solves = b->d_zsolve;
        z.nitems = c.nitems;
        z.nz = 16 * c.key;  // selects the launch template; per-member lengths come from solves
        if (f32) {
            z.tau = b->tau;
            z.pinned_zero = 1;
            PPB_CHECK(launch_zsolve_pcr_f32(16 * c.key, z, b->stream));
        } else {
            PPB_CHECK(launch_zsolve_pcr_f64(16 * c.key, z, b->stream));
        }
    }
    return PP_SUCCESS;
}

int run_fp64(pp_batch_s *b)
{
    int rc;
    if ((rc = run_line_stage(b, b->lx, FORM_F64, 0, 1, b->d_rhs, b->d_sol, BF_RHS, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = run_line_stage(b, b->ly, FORM_F64, 1, 1, b->d_sol, b->d_sol, 0, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = run_z_stage(b, false))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = run_line_stage(b, b->ly, FORM_F64, 1, 0, b->d_sol, b->d_sol, 0, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = run_line_stage(b, b->lx, FORM_F64, 0, 0, b->d_sol, b->d_sol, BF_SCALE, nullptr, nullptr, 0))) return rc;
    return mark(b);
}

// y forward, z (FP32 + FP64 near-singular columns), y inverse on the work arena
int mixed_inner(pp_batch_s *b)
{
    int rc;
    if ((rc = run_line_stage(b, b->ly32, FORM_F32, 1, 1, b->d_w, b->d_w, 0, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = run_z_stage(b, true))) return rc;
    if (b->nfix) {
        ZArgs z{};
        z.data = b->d_w;
        z.solves = b->d_zsolve;
        PPB_CHECK(launch_zfix64_f32(z, b->d_fix, b->nfix, b->nz_max, b->stream));
    }
    if ((rc = mark(b))) return rc;
    if ((rc = run_line_stage(b, b->ly32, FORM_F32, 1, 0, b->d_w, b->d_w, 0, nullptr, nullptr, 0))) return rc;
    return mark(b);
}

// the mixed form over the batch (same steps as the single-solve handle's)
int run_mixed(pp_batch_s *b)
{
    int rc;
    double *sums = b->nm00 ? b->d_sums : nullptr;
    const double *plane = b->nm00 ? b->d_plane : nullptr;
    if ((rc = run_line_stage(b, b->lx32, FORM_F64_TO_F32, 0, 1, b->d_rhs, b->d_w, BF_RHS, sums, nullptr, 0))) return rc;
    if ((rc = mark(b))) return rc;
    if ((rc = mixed_inner(b))) return rc;
    if (b->nm00) PPB_CHECK(launch_mode00_batch(b->d_m00, b->nm00, b->nz_max, 0, b->stream));
    if ((rc = run_line_stage(b, b->lx32, FORM_F32_TO_F64, 0, 0, b->d_w, b->d_sol, BF_SCALE, nullptr, plane, 0)))
        return rc;
    if ((rc = mark(b))) return rc;
    for (int it = 0; it < b->refine; it++) {
        PPB_CHECK(launch_residual_f32_batch(b->d_res, b->d_res_items, b->nres_items, b->stream));
        if ((rc = mark(b))) return rc;
        if ((rc = run_line_stage(b, b->lx32, FORM_F32, 0, 1, b->d_w, b->d_w, 0, sums, nullptr, 0))) return rc;
        if ((rc = mark(b))) return rc;
        if ((rc = mixed_inner(b))) return rc;
        if (b->nm00) PPB_CHECK(launch_mode00_batch(b->d_m00, b->nm00, b->nz_max, 1, b->stream));
        if ((rc = run_line_stage(b, b->lx32, FORM_F32_TO_F64, 0, 0, b->d_w, b->d_sol, BF_SCALE, nullptr, plane, 1)))
            return rc;
        if ((rc = mark(b))) return rc;
    }
    return PP_SUCCESS;
}

// mixed-form tables: FP32 launch classes, work arena, pinned members, near-singular
// columns (with the kernels' eigenvalue rounding), residual descriptors
int build_mixed(pp_batch_s *b)
{
    int rc;
    if ((rc = build_line_classes<float>(b, 0, b->lx32)) || (rc = build_line_classes<float>(b, 1, b->ly32))) return rc;
    if (hipMalloc((void **)&b->d_w, (size_t)b->arena_elems * sizeof(float)) != hipSuccess) return PP_ALLOCATION_FAIL;
    long so = 0, po = 0, uo = 0;
    std::vector<int4> fix;
    std::vector<Mode00Item> m00;
    std::vector<ResArgs> res(b->mem.size());
    std::vector<int4> ritems;
    for (size_t i = 0; i < b->mem.size(); i++) {
        BatchMember &m = b->mem[i];
        const SolveSetup &g = m.g;
        m.sums_off = so;
        so += (long)g.ny * g.nz;
        m.plane_off = po;
        po += g.nz + 2;
        m.upart_off = uo;
        uo += residual_parts(g.nx, g.ny);
        const double Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
        const double Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
        const std::vector<double> lx = eig_factors(g, 0), ly = eig_factors(g, 1);
        m.pinned = false;
        for (int sy = 0; sy < g.ny; sy++)
            for (int sx = 0; sx < g.nx; sx++) {
                const double eig = column_eig(lx, ly, sx, sy, Cx, Cy);
                bool pin = std::fabs(eig + 2.) < 1.e-10;
                if (pin && b->zero_mode == 1 && (g.zlo_d || g.zhi_d)) pin = false;
                if (pin) {
                    m.pinned = true;
                    continue;
                }
                if (std::fabs(eig + 2.) < b->tau) fix.push_back(make_int4(sx, sy, (int)i, 0));
            }
    }
    const long nsums = so, nplane = po, nupart = uo;
    if (hipMalloc((void **)&b->d_sums, (size_t)nsums * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&b->d_plane, (size_t)nplane * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&b->d_upart, (size_t)nupart * sizeof(double)) != hipSuccess)
        return PP_ALLOCATION_FAIL;
    // members without a pinned column add zeros in the inverse-x store
    PPB_CHECK(hipMemset(b->d_plane, 0, (size_t)nplane * sizeof(double)));
    for (size_t i = 0; i < b->mem.size(); i++) {
        const BatchMember &m = b->mem[i];
        const SolveSetup &g = m.g;
        if (m.pinned) {
            Mode00Item it{};
            it.line_sums = b->d_sums + m.sums_off;
            it.ny = g.ny;
            it.nz = g.nz;
            it.b0 = g.zlo_d ? -2.0 - 1. : -2.0 + 1.;
            it.bn = g.zhi_d ? -2.0 - 1. : -2.0 + 1.;
            it.inv_scale = 1. / g.scale;
            it.plane_add = b->d_plane + m.plane_off;
            it.upart = b->d_upart + m.upart_off;
            it.nparts = residual_parts(g.nx, g.ny);
            m00.push_back(it);
        }
        ResArgs &r = res[i];
        r.f = b->d_rhs + m.offset;
        r.u = b->d_sol + m.offset;
        r.r = b->d_w + m.offset;
        r.nx = g.nx;
        r.ny = g.ny;
        r.nz = g.nz;
        r.c3 = g.dxyz[2];
        r.Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
        r.Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
        for (int f = 0; f < 6; f++) r.bc[f] = g.bc[f];
        r.dtab = b->d_dtab + m.dtab_off;
        r.dmask = g.dmask;
        r.pin_defect = m.pinned ? b->d_plane + m.plane_off + g.nz : nullptr;
        r.upart = m.pinned ? b->d_upart + m.upart_off : nullptr;
        const int gx = (g.nx + 63) / 64, gy = (g.ny + 3) / 4, gz = (g.nz + 7) / 8;
        if ((long)gx * gy * gz != residual_blocks(g.nx, g.ny, g.nz)) return PP_INPUT_ARGS;
        for (int bz = 0; bz < gz; bz++)
            for (int by = 0; by < gy; by++)
                for (int bx = 0; bx < gx; bx++) ritems.push_back(make_int4((int)i, bx, by, bz));
    }
    b->nfix = (int)fix.size();
    b->nm00 = (int)m00.size();
    b->nres_items = (int)ritems.size();
    if ((rc = upload_vec(&b->d_fix, fix)) || (rc = upload_vec(&b->d_m00, m00)) || (rc = upload_vec(&b->d_res, res)) ||
        (rc = upload_vec(&b->d_res_items, ritems)))
        return rc;
    return PP_SUCCESS;
}

}  // namespace

extern "C" {

int pp_batch_create(pp_batch *out)
{
    if (!out) return PP_INPUT_ARGS;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PP_NO_DEVICE;
    pp_batch_s *b = new (std::nothrow) pp_batch_s();
    if (!b) return PP_ALLOCATION_FAIL;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return PP_HIP_ERROR;
    }
    b->own_stream = true;
    *out = b;
    return PP_SUCCESS;
}

int pp_batch_destroy(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_device(b);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PP_SUCCESS;
}

int pp_batch_add(pp_batch b, int nx, int ny, int nz, const char *bc, const double *X, int *id)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!bc || !X || !id) return PP_INPUT_ARGS;
    if (b->finalized) return PP_CALL_ORDER;
    const bool zok = b->precision == 1 ? zsolve_supported_f32(nz) : zsolve_supported_f64(nz);
    if (!line_size_supported(nx) || !line_size_supported(ny) || !zok) return PP_UNSUPPORTED_SIZE;
    if (!(X[1] > X[0] && X[3] > X[2] && X[5] > X[4])) return PP_INPUT_ARGS;
    if (16.0 * nx * ny * sizeof(double) >= 4294967296.0) return PP_UNSUPPORTED_SIZE;
    BatchMember m;
    m.g.nx = nx;
    m.g.ny = ny;
    m.g.nz = nz;
    const int rc = setup_boundary(m.g, bc);
    if (rc) return rc;
    for (int i = 0; i < 6; i++) m.g.X[i] = X[i];
    setup_coords(m.g);
    *id = (int)b->mem.size();
    b->mem.push_back(m);
    return PP_SUCCESS;
}

int pp_batch_finalize(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    if (b->finalized) return PP_CALL_ORDER;
    if (b->mem.empty()) return PP_INPUT_ARGS;
    // arenas, eigenvalue factors and Dirichlet tables, packed per member
    long off = 0, loff = 0, doff = 0;
    std::vector<double> lam, dt;
    b->nz_max = 0;
    for (BatchMember &m : b->mem) {
        m.offset = off;
        off += ((long)m.g.nx * m.g.ny * m.g.nz + 63) / 64 * 64;  // 512-byte aligned fields
        b->nz_max = std::max(b->nz_max, m.g.nz);
        for (int a = 0; a < 2; a++) {
            m.lam_off[a] = loff;
            const std::vector<double> l = eig_factors(m.g, a);
            lam.insert(lam.end(), l.begin(), l.end());
            loff += (long)l.size();
        }
        m.dtab_off = doff;
        const std::vector<double> d = dirichlet_table(m.g);
        dt.insert(dt.end(), d.begin(), d.end());
        doff += (long)d.size();
    }
    b->arena_elems = off;
    int rc;
    if (hipMalloc((void **)&b->d_rhs, (size_t)off * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&b->d_sol, (size_t)off * sizeof(double)) != hipSuccess) {
        free_device(b);
        return PP_ALLOCATION_FAIL;
    }
    if ((rc = upload_vec(&b->d_lam, lam)) || (rc = upload_vec(&b->d_dtab, dt))) {
        free_device(b);
        return rc;
    }
    if (b->precision == 1 && (rc = build_mixed(b))) {
        free_device(b);
        return rc;
    }
    std::vector<LineSolve> ls(b->mem.size());
    std::vector<ZSolve> zs(b->mem.size());
    for (size_t i = 0; i < b->mem.size(); i++) {
        const BatchMember &m = b->mem[i];
        LineSolve &l = ls[i];
        l.offset = m.offset;
        l.nx = m.g.nx;
        l.ny = m.g.ny;
        l.nz = m.g.nz;
        l.dmask = m.g.dmask;
        l.dz = m.g.dxyz[2];
        l.inv_scale = 1. / m.g.scale;
        l.dtab = b->d_dtab + m.dtab_off;
        l.sums_off = m.sums_off;
        l.plane_off = m.plane_off;
        ZSolve &z = zs[i];
        z.offset = m.offset;
        z.SX = m.g.nx;
        z.SY = m.g.ny;
        z.nz = m.g.nz;
        z.zlo_dirichlet = m.g.zlo_d;
        z.zhi_dirichlet = m.g.zhi_d;
        z.Cx = (m.g.dxyz[2] / m.g.dxyz[0]) * (m.g.dxyz[2] / m.g.dxyz[0]);
        z.Cy = (m.g.dxyz[2] / m.g.dxyz[1]) * (m.g.dxyz[2] / m.g.dxyz[1]);
        z.lamx = b->d_lam + m.lam_off[0];
        z.lamy = b->d_lam + m.lam_off[1];
    }
    if ((rc = upload_vec(&b->d_lsolve, ls)) || (rc = upload_vec(&b->d_zsolve, zs)) || (rc = build_z_classes(b))) {
        free_device(b);
        return rc;
    }
    if (b->precision == 0 &&
        ((rc = build_line_classes<double>(b, 0, b->lx)) || (rc = build_line_classes<double>(b, 1, b->ly)))) {
        free_device(b);
        return rc;
    }
    b->finalized = true;
    return PP_SUCCESS;
}

int pp_batch_count(pp_batch b, int *members, long long *points)
{
    if (!b) return PP_INVALID_HANDLE;
    long long p = 0;
    for (const BatchMember &m : b->mem) p += (long long)m.g.nx * m.g.ny * m.g.nz;
    if (members) *members = (int)b->mem.size();
    if (points) *points = p;
    return PP_SUCCESS;
}

int pp_batch_launch_count(pp_batch b, int *per_stage)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (!per_stage) return PP_INPUT_ARGS;
    per_stage[0] = (int)(b->precision ? b->lx32.size() : b->lx.size());
    per_stage[1] = (int)(b->precision ? b->ly32.size() : b->ly.size());
    per_stage[2] = (int)b->lz.size();
    return PP_SUCCESS;
}

int pp_batch_field(pp_batch b, int id, double **d_rhs)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (id < 0 || id >= (int)b->mem.size() || !d_rhs) return PP_INPUT_ARGS;
    *d_rhs = b->d_rhs + b->mem[id].offset;
    return PP_SUCCESS;
}

int pp_batch_solution(pp_batch b, int id, double **d_u)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (id < 0 || id >= (int)b->mem.size() || !d_u) return PP_INPUT_ARGS;
    *d_u = b->d_sol + b->mem[id].offset;
    return PP_SUCCESS;
}

int pp_batch_set_rhs(pp_batch b, int id, const double *rhs)
{
    double *d = nullptr;
    int rc = pp_batch_field(b, id, &d);
    if (rc) return rc;
    if (!rhs) return PP_INPUT_ARGS;
    const BatchMember &m = b->mem[id];
    PPB_CHECK(hipMemcpyAsync(d, rhs, (size_t)m.g.nx * m.g.ny * m.g.nz * sizeof(double), hipMemcpyHostToDevice,
                             b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_get_solution(pp_batch b, int id, double *u)
{
    double *d = nullptr;
    int rc = pp_batch_solution(b, id, &d);
    if (rc) return rc;
    if (!u) return PP_INPUT_ARGS;
    const BatchMember &m = b->mem[id];
    PPB_CHECK(hipMemcpyAsync(u, d, (size_t)m.g.nx * m.g.ny * m.g.nz * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_get_rhs(pp_batch b, int id, double *rhs)
{
    double *d = nullptr;
    int rc = pp_batch_field(b, id, &d);
    if (rc) return rc;
    if (!rhs) return PP_INPUT_ARGS;
    const BatchMember &m = b->mem[id];
    PPB_CHECK(hipMemcpyAsync(rhs, d, (size_t)m.g.nx * m.g.ny * m.g.nz * sizeof(double), hipMemcpyDeviceToHost,
                             b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_fill_uniform(pp_batch b, unsigned long long seed, int round_f32)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    for (size_t i = 0; i < b->mem.size(); i++) {
        const BatchMember &m = b->mem[i];
        PPB_CHECK(launch_fill_uniform(b->d_rhs + m.offset, (long)m.g.nx * m.g.ny * m.g.nz, seed + i, round_f32,
                                      b->stream));
    }
    return PP_SUCCESS;
}

int pp_batch_set_stream(pp_batch b, void *stream)
{
    if (!b) return PP_INVALID_HANDLE;
    PPB_CHECK(hipStreamSynchronize(b->stream));
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    b->own_stream = false;
    b->stream = (hipStream_t)stream;
    return PP_SUCCESS;
}

int pp_batch_set_option(pp_batch b, int key, long long value)
{
    if (!b) return PP_INVALID_HANDLE;
    switch (key) {
    case PP_OPT_ZERO_MODE:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        if (b->finalized) return PP_CALL_ORDER;
        b->zero_mode = (int)value;
        return PP_SUCCESS;
    case PP_OPT_PRECISION:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        if (b->finalized || !b->mem.empty()) return PP_CALL_ORDER;
        b->precision = (int)value;
        return PP_SUCCESS;
    case PP_OPT_REFINE:
        if (value < 0 || value > 16) return PP_INPUT_ARGS;
        b->refine = (int)value;
        return PP_SUCCESS;
    case PP_OPT_TIMING: {
        if (value < 0 || value > 1024) return PP_INPUT_ARGS;
        PPB_CHECK(hipStreamSynchronize(b->stream));
        for (auto &e : b->ev)
            if (e) (void)hipEventDestroy(e);
        b->ev.assign((size_t)pp_batch_s::kMarks * value, nullptr);
        for (auto &e : b->ev)
            if (hipEventCreate(&e) != hipSuccess) return PP_HIP_ERROR;
        b->ev_slots = (int)value;
        b->ev_next = b->ev_used = 0;
        b->timing = value != 0;
        return PP_SUCCESS;
    }
    default: return PP_INPUT_ARGS;
    }
}

int pp_batch_set_option_double(pp_batch b, int key, double value)
{
    if (!b) return PP_INVALID_HANDLE;
    if (key != PP_OPT_TAU || !(value >= 0.0 && value < 4.0)) return PP_INPUT_ARGS;
    if (b->finalized) return PP_CALL_ORDER;
    b->tau = value;
    return PP_SUCCESS;
}

// solve every member: right-hand sides kept, solutions in the solution arena; each stage is
// one launch per shape class
int pp_batch_solve(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    b->ev_cur = nullptr;
    if (b->timing) {
        b->ev_cur = &b->ev[(size_t)pp_batch_s::kMarks * b->ev_next];
        b->ev_next = (b->ev_next + 1) % b->ev_slots;
        if (b->ev_used < b->ev_slots) b->ev_used++;
        b->marks_cur = 0;
    }
    int rc = mark(b);
    if (rc) return rc;
    rc = b->precision == 1 ? run_mixed(b) : run_fp64(b);
    if (b->ev_cur) b->marks_last = b->marks_cur;
    b->ev_cur = nullptr;
    return rc;
}

int pp_batch_get_kernel_times(pp_batch b, float *ms, int max, int *count)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!ms || !count || max < 0) return PP_INPUT_ARGS;
    if (!b->timing || b->ev_used == 0 || b->marks_last < 2) return PP_CALL_ORDER;
    PPB_CHECK(hipStreamSynchronize(b->stream));
    const int n = b->marks_last - 1;
    std::vector<double> acc(n, 0.0);
    for (int s = 0; s < b->ev_used; s++) {
        hipEvent_t *ev = &b->ev[(size_t)pp_batch_s::kMarks * s];
        for (int i = 0; i < n; i++) {
            float t = 0.f;
            PPB_CHECK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
            acc[i] += t;
        }
    }
    *count = n;
    for (int i = 0; i < n && i < max; i++) ms[i] = (float)(acc[i] / b->ev_used);
    b->ev_used = b->ev_next = 0;
    return PP_SUCCESS;
}

int pp_batch_synchronize(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

}  // extern "C"
