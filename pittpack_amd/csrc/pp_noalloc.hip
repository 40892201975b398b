// pp_noalloc.hip -- pp_solve_device_oop_noalloc: the FP64 solve of pp_solve_device_oop for
// callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  No handle: the sizes, boundary conditions, box and
// zero-mode rule come as arguments, and every table the kernels read (twiddles, eigenvalue
// factors, Dirichlet term, periodic-z tables) lives in one file-scope device array `scratch`
// of kScratchDoubles doubles (2 MiB), which each GPU the library runs on holds until the
// process ends.  The z solve works on the solution array in place (no z-tile buffer).
//
// Every call on a GPU shares that GPU's array and leaves its tables in it; a call whose
// sizes, boundary conditions, box and zero-mode rule equal the previous call's on that GPU
// reuses them and only enqueues kernels (asynchronous, capturable), any other call rebuilds
// them on the host (static staging), copies them in and waits for that copy.  Calls that
// overlap on one GPU race (on the tables and on this file's host state): serialise them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "pittpack_hip.h"
#include "pp_handle.h"
#include "pp_launch.h"
#include "pp_setup.h"

namespace {

using namespace pp;

constexpr size_t kScratchDoubles = 256 * 1024;  // 2 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

// the tables of the last call per device (host side) and where they sit in scratch
struct Resident {
    bool valid = false;
    int nx = 0, ny = 0, nz = 0, zero_mode = 0;
    char bc[6] = {};
    double X[6] = {};
    size_t tab[2][5] = {};  // offsets (doubles) of tw, A, B, C, D per axis
    size_t ztab[5] = {};    // periodic z without a compiled complex plan: real z-line tables
    size_t lamx = 0, lamy = 0, dtab = 0, ztw = 0, zc2 = 0;
    int zsing_x = -1, zsing_y = -1;
    double zsing_eig = -2.0;
    size_t used = 0;  // doubles of scratch in use
};

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
Resident g_res[kMaxDevices];
double g_host[kScratchDoubles];       // host staging of the tables

bool same_key(const Resident &r, int nx, int ny, int nz, const char *bc, const double *X, int zero_mode)
{
    return r.valid && r.nx == nx && r.ny == ny && r.nz == nz && r.zero_mode == zero_mode &&
           std::memcmp(r.bc, bc, 6) == 0 && std::memcmp(r.X, X, sizeof(r.X)) == 0;
}

// lay the tables of `g` out in g_host (offsets into r); false if they exceed kScratchDoubles
bool build_host_tables(const SolveSetup &g, Resident &r)
{
    size_t o = 0;
    auto take = [&](size_t n) {
        const size_t at = o;
        o += (n + 31) / 32 * 32;  // 256-byte aligned sub-arrays
        return at;
    };
    const int n[2] = {g.nx, g.ny};
    for (int a = 0; a < 2; a++) {
        const AxisTableSizes z = axis_table_sizes(n[a]);
        for (int i = 0; i < 5; i++) r.tab[a][i] = take(z.n[i]);
    }
    r.lamx = take((size_t)g.nx);
    r.lamy = take((size_t)g.ny);
    r.dtab = take(dirichlet_table_size(g));
    const bool zc = g.zper && zperiodic_supported(g.nz);
    AxisTableSizes zs{};
    if (g.zper) {
        zs = axis_table_sizes(zc ? 2 * g.nz : g.nz);
        if (zc) {
            r.ztw = take(zs.n[0]);
        } else {
            for (int i = 0; i < 5; i++) r.ztab[i] = take(zs.n[i]);
        }
        r.zc2 = take((size_t)g.nz);
    }
    if (o > kScratchDoubles) return false;
    std::memset(g_host, 0, o * sizeof(double));
    for (int a = 0; a < 2; a++)
        fill_axis_tables(n[a], g_host + r.tab[a][0], g_host + r.tab[a][1], g_host + r.tab[a][2], g_host + r.tab[a][3],
                         g_host + r.tab[a][4]);
    fill_eig_factors(g, 0, g_host + r.lamx);
    fill_eig_factors(g, 1, g_host + r.lamy);
    fill_dirichlet_table(g, g_host + r.dtab);
    r.zsing_x = r.zsing_y = -1;
    if (g.zper) {
        if (zc) {
            // complex FFT of length nz: only its twiddles are read
            static double unused[4][2 * 1024 + 2];
            fill_axis_tables(2 * g.nz, g_host + r.ztw, unused[0], unused[1], unused[2], unused[3]);
        } else {
            fill_axis_tables(g.nz, g_host + r.ztab[0], g_host + r.ztab[1], g_host + r.ztab[2], g_host + r.ztab[3],
                             g_host + r.ztab[4]);
        }
        fill_zper_cos2(g.nz, g_host + r.zc2);
        // the singular column, with the kernels' eigenvalue rounding (no contraction)
        const double Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
        const double Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
        const double *lx = g_host + r.lamx, *ly = g_host + r.lamy;
        for (int sy = 0; sy < g.ny && r.zsing_x < 0; sy++)
            for (int sx = 0; sx < g.nx; sx++) {
                volatile double t1 = lx[sx] * Cx;
                volatile double t2 = -2.0 + t1;
                volatile double t3 = ly[sy] * Cy;
                const double eig = t2 + t3;
                if (std::fabs(eig + 2.) < 1.e-10) {
                    r.zsing_x = sx;
                    r.zsing_y = sy;
                    r.zsing_eig = eig;
                    break;
                }
            }
    }
    r.used = o;
    return true;
}

}  // namespace

extern "C" int pp_solve_device_oop_noalloc(int nx, int ny, int nz, const char *bc, const double *X, int zero_mode,
                                           const double *d_rhs, double *d_u, void *stream)
{
    if (!bc || !X || !d_rhs || !d_u) return PP_INPUT_ARGS;
    if (zero_mode != 0 && zero_mode != 1) return PP_INPUT_ARGS;
    if (nx < 2 || ny < 2 || nz < 4) return PP_INPUT_ARGS;
    if (!(X[1] > X[0] && X[3] > X[2] && X[5] > X[4])) return PP_INPUT_ARGS;
    if (!line_size_supported(nx) || !line_size_supported(ny)) return PP_UNSUPPORTED_SIZE;
    SolveSetup g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    int rc = setup_boundary(g, bc);
    if (rc) return rc;
    for (int i = 0; i < 6; i++) g.X[i] = X[i];
    setup_coords(g);
    if (g.zper) {
        if (!zperiodic_supported(nz) && !line_size_supported(nz)) return PP_UNSUPPORTED_SIZE;
    } else if (!zsolve_supported_f64(nz) && !zsolve_long_supported_f64(nz)) {
        return PP_UNSUPPORTED_SIZE;
    }
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PP_NO_DEVICE;
    if (dev < 0 || dev >= kMaxDevices) return PP_UNSUPPORTED_SIZE;

    Resident r;
    double *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            // once per device, with that device current: the array's device address (its
            // host-side name is not a device pointer)
            void *p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PP_HIP_ERROR;
            g_scratch[dev] = static_cast<double *>(p);
        }
        base = g_scratch[dev];
        if (!same_key(g_res[dev], nx, ny, nz, bc, X, zero_mode)) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(s, &cs) != hipSuccess) return PP_HIP_ERROR;
            // new tables cannot be copied in under capture: one uncaptured call with the same
            // arguments makes them resident first
            if (cs != hipStreamCaptureStatusNone) return PP_CALL_ORDER;
            g_res[dev].valid = false;
            Resident nr;
            if (!build_host_tables(g, nr)) return PP_UNSUPPORTED_SIZE;
            if (hipMemcpyAsync(base, g_host, nr.used * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return PP_HIP_ERROR;
            nr.nx = nx;
            nr.ny = ny;
            nr.nz = nz;
            nr.zero_mode = zero_mode;
            std::memcpy(nr.bc, bc, 6);
            std::memcpy(nr.X, X, sizeof(nr.X));
            nr.valid = true;
            g_res[dev] = nr;
        }
        r = g_res[dev];
    }

    // a handle view of the solve: tables in scratch, the in-place z layout, default shapes
    pp_solver h;
    h.g = g;
    h.bc_set = h.box_set = h.coords_set = true;
    h.zero_mode = zero_mode;
    h.zsolver = 0;
    h.ztiles = 0;
    h.zthomas_tau = 0.0;  // no Thomas columns: their list and factors would need allocations
    h.stream = s;
    for (int a = 0; a < 2; a++) {
        void **t[5] = {&h.tab64[a].tw, &h.tab64[a].A, &h.tab64[a].B, &h.tab64[a].C, &h.tab64[a].D};
        for (int i = 0; i < 5; i++) *t[i] = base + r.tab[a][i];
    }
    h.d_lamx = base + r.lamx;
    h.d_lamy = base + r.lamy;
    h.d_dtab = base + r.dtab;
    if (g.zper) {
        if (zperiodic_supported(nz)) {
            h.d_ztw = base + r.ztw;
        } else {
            void **t[5] = {&h.tabz.tw, &h.tabz.A, &h.tabz.B, &h.tabz.C, &h.tabz.D};
            for (int i = 0; i < 5; i++) *t[i] = base + r.ztab[i];
        }
        h.d_zc2 = base + r.zc2;
        h.zsing_x = r.zsing_x;
        h.zsing_y = r.zsing_y;
        h.zsing_eig = r.zsing_eig;
    }
    h.tables_ready = true;
    return run_fp64(&h, d_rhs, d_u);
}
