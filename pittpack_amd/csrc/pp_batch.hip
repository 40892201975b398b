// pp_batch.hip -- batched form: many independent solves of uneven sizes and boundary
// conditions, each stage run as one launch per shape class over all members of the class.
//
// The reference solves one grid per PencilDcmp object (src/pencilDcmp.cpp:5801-6430); here
// the members' fields sit back to back in device arenas and every kernel launch covers the
// blocks of all members that share its template (line length and transform kind for the x
// and y stages, rows per lane for the z stage).  A block finds its member through a
// per-launch item table; per-member scalars (sizes, dz, 1/(NxNy), Dirichlet tables,
// eigenvalue factors) come from a per-member descriptor.  Kernels and arithmetic are those
// of the single-solve path, so a member's result does not depend on the batch around it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <tuple>
#include <vector>

#include "pittpack_hip.h"
#include "pp_batch_core.h"
#include "pp_launch.h"
#include "pp_line_bs.h"
#include "pp_resx.h"

#define PPB_CHECK(x)                                  \
    do {                                              \
        if ((x) != hipSuccess) return PP_HIP_ERROR;   \
    } while (0)

namespace pp {

namespace {

size_t take(size_t &cursor, size_t bytes)
{
    const size_t o = (cursor + 255) / 256 * 256;
    cursor = o + bytes;
    return o;
}

int mark(hipStream_t s, Marks *mk)
{
    if (mk && mk->ev && mk->n < mk->max) PPB_CHECK(hipEventRecord(mk->ev[mk->n++], s));
    return PP_SUCCESS;
}

template <typename T>
int put(T *dst, const std::vector<T> &v, hipStream_t s)
{
    if (v.empty()) return PP_SUCCESS;
    PPB_CHECK(hipMemcpyAsync(dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return PP_SUCCESS;
}

}  // namespace

TableCache::~TableCache()
{
    for (auto &kv : t_)
        for (double *p : {kv.second.tw, kv.second.A, kv.second.B, kv.second.C, kv.second.D})
            if (p) (void)hipFree(p);
}

int TableCache::get(int N, AxisDev &out)
{
    auto it = t_.find(N);
    if (it != t_.end()) {
        out = it->second;
        return PP_SUCCESS;
    }
    const AxisTables &t = axis_tables_cached(N);
    AxisDev d;
    auto up = [](double **dst, const std::vector<double> &v) -> int {
        if (hipMalloc((void **)dst, v.size() * sizeof(double)) != hipSuccess) return PP_ALLOCATION_FAIL;
        if (hipMemcpy(*dst, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return PP_HIP_ERROR;
        return PP_SUCCESS;
    };
    int rc;
    if ((rc = up(&d.tw, t.tw)) || (rc = up(&d.A, t.A)) || (rc = up(&d.B, t.B)) || (rc = up(&d.C, t.C)) ||
        (rc = up(&d.D, t.D)))
        return rc;
    t_[N] = d;
    out = d;
    return PP_SUCCESS;
}

int BatchCore::add(int nx, int ny, int nz, const char *bc, const double *X, int *id)
{
    if (!bc || !X || !id) return PP_INPUT_ARGS;
    if (planned) return PP_CALL_ORDER;
    BatchMember m;
    m.g.nx = nx;
    m.g.ny = ny;
    m.g.nz = nz;
    const bool zper = bc[4] == 'P';
    // periodic z: any length the single-solve handle takes (the mixed form from 3 rows);
    // otherwise FP64 columns up to 1024 rows (513..1024 by the long-column kernel, one launch
    // per member) and FP32 columns up to 1024
    bool zok;
    if (zper) zok = (zperiodic_supported(nz) || line_size_supported(nz)) && (precision == 0 || nz >= 3);
    else zok = precision == 1 ? zsolve_supported_f32(nz) : (zsolve_supported_f64(nz) || zsolve_long_supported_f64(nz));
    if (!line_size_supported(nx) || !line_size_supported(ny) || !zok) return PP_UNSUPPORTED_SIZE;
    if (!(X[1] > X[0] && X[3] > X[2] && X[5] > X[4])) return PP_INPUT_ARGS;
    if (16.0 * nx * ny * sizeof(double) >= 4294967296.0) return PP_UNSUPPORTED_SIZE;
    // FP64 members hand the z solve 16-column tiles, which y blocks of 16 columns fill alone
    // and narrower y blocks (a divisor of 16) fill together; periodic z stays in place
    if (precision == 0 && !zper) {
        const int ly = line_lines_any<double>(ny, 1);
        if (ly <= 0 || zsolve_cpb<double>() % ly != 0) return PP_UNSUPPORTED_SIZE;
    }
    const int rc = setup_boundary(m.g, bc);
    if (rc) return rc;
    for (int i = 0; i < 6; i++) m.g.X[i] = X[i];
    setup_coords(m.g);
    *id = (int)mem.size();
    mem.push_back(m);
    return PP_SUCCESS;
}

long long BatchCore::points() const
{
    long long p = 0;
    for (const BatchMember &m : mem) p += (long long)m.g.nx * m.g.ny * m.g.nz;
    return p;
}

// x (axis 0) or y (axis 1) stage: one class per (N, kind); block sizes of precision R
template <typename R>
int BatchCore::plan_lines(int axis, std::vector<HostClass> &out)
{
    std::map<std::pair<int, int>, std::vector<LineItem>> cls;
    for (int m = 0; m < (int)mem.size(); m++) {
        const SolveSetup &g = mem[m].g;
        const int N = axis == 0 ? g.nx : g.ny;
        const int lines = line_lines_any<R>(N, axis);
        if (lines <= 0) return PP_UNSUPPORTED_SIZE;
        auto &items = cls[{N, g.kind[axis]}];
        if (axis == 0) {
            const long nl = (long)g.ny * g.nz;
            for (long bx = 0; bx < (nl + lines - 1) / lines; bx++) items.push_back({m, (int)bx, 0});
        } else {
            const int nbx = (g.nx + lines - 1) / lines;
            for (int z = 0; z < g.nz; z++)
                for (int bx = 0; bx < nbx; bx++) items.push_back({m, bx, z});
        }
    }
    out.clear();
    for (auto &kv : cls) {
        HostClass c;
        c.key = kv.first.first;
        c.kind = kv.first.second;
        c.nitems = (int)kv.second.size();
        c.li = std::move(kv.second);
        if (sizeof(R) == 8 && axis == 0 && line_has_variant(c.key)) {
            // x blocks of the alternative shape hold a different number of lines
            const int la = line_block_lines_alt(c.key, 0);
            for (int m = 0; m < (int)mem.size(); m++) {
                const SolveSetup &g = mem[m].g;
                if (g.nx != c.key || g.kind[0] != c.kind) continue;
                const long nl = (long)g.ny * g.nz;
                for (long bx = 0; bx < (nl + la - 1) / la; bx++) c.li_alt.push_back({m, (int)bx, 0});
            }
        }
        out.push_back(std::move(c));
    }
    return PP_SUCCESS;
}

void BatchCore::plan_z()
{
    std::map<int, std::vector<ZItem>> cls, pcls;
    const int cpb = precision == 1 ? zsolve_cpb<float>() : zsolve_cpb<double>();
    zpl_.clear();
    zlong_.clear();
    zth_.clear();
    zth_tab_.clear();
    smt_.clear();
    zc2_.clear();
    zc2_off_.clear();
    smt_nz_max_ = 0;
    for (int m = 0; m < (int)mem.size(); m++) {
        const SolveSetup &g = mem[m].g;
        if (g.zper) {
            if (!zc2_off_.count(g.nz)) {
                zc2_off_[g.nz] = (long)zc2_.size();
                const std::vector<double> c2 = zper_cos2(g.nz);
                zc2_.insert(zc2_.end(), c2.begin(), c2.end());
            }
            if (zperiodic_supported(g.nz)) {
                auto &items = pcls[g.nz];
                const int lines = zper_lines(g.nz);
                const int nbx = (g.nx + lines - 1) / lines;
                for (int sy = 0; sy < g.ny; sy++)
                    for (int bx = 0; bx < nbx; bx++) items.push_back({m, bx, sy});
            } else {
                zpl_.push_back(m);
            }
            // FP64: the singular column by the reference's branch (the mixed form solves it
            // from plane sums)
            if (precision == 0) {
                const std::vector<double> lx = eig_factors(g, 0), ly = eig_factors(g, 1);
                const double Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
                const double Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
                bool found = false;
                for (int sy = 0; sy < g.ny && !found; sy++)
                    for (int sx = 0; sx < g.nx && !found; sx++) {
                        const double eig = column_eig(lx, ly, sx, sy, Cx, Cy);
                        if (std::fabs(eig + 2.) < 1.e-10) {
                            smt_.push_back({mem[m].offset, g.nx, g.ny, g.nz, sx, sy, eig});
                            smt_nz_max_ = std::max(smt_nz_max_, g.nz);
                            found = true;
                        }
                    }
            }
            continue;
        }
        if (precision == 0 && zsolve_long_supported_f64(g.nz)) {
            zlong_.push_back(m);
            continue;
        }
        if (precision == 0 && zthomas_tau > 0.0) {
            // nearly singular columns, as the single-solve handle lists them
            const std::vector<double> lx = eig_factors(g, 0), ly = eig_factors(g, 1);
            const double Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
            const double Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
            for (int sy = 0; sy < g.ny; sy++)
                for (int sx = 0; sx < g.nx; sx++) {
                    const double eig = column_eig(lx, ly, sx, sy, Cx, Cy);
                    bool pin = std::fabs(eig + 2.) < 1.e-10;
                    if (pin && zero_mode == 1 && (g.zlo_d || g.zhi_d)) pin = false;
                    if (pin || !(std::fabs(eig + 2.) < zthomas_tau)) continue;
                    zth_.push_back({sx, sy, m, 0, (long)zth_tab_.size()});
                    zth_tab_.resize(zth_tab_.size() + 2 * (size_t)g.nz);
                    zthomas_factors(eig, g.zlo_d ? eig - 1. : eig + 1., g.zhi_d ? eig - 1. : eig + 1., g.nz,
                                    zth_tab_.data() + zth_tab_.size() - 2 * (size_t)g.nz);
                }
        }
        auto &items = cls[(g.nz + 15) / 16];
        const int nbx = (g.nx + cpb - 1) / cpb;
        for (int sy = 0; sy < g.ny; sy++)
            for (int bx = 0; bx < nbx; bx++) items.push_back({m, bx, sy});
    }
    lz_.clear();
    for (auto &kv : cls) {
        HostClass c;
        c.key = kv.first;
        c.nitems = (int)kv.second.size();
        c.zi = std::move(kv.second);
        lz_.push_back(std::move(c));
    }
    lzp_.clear();
    for (auto &kv : pcls) {
        HostClass c;
        c.key = kv.first;
        c.nitems = (int)kv.second.size();
        c.zi = std::move(kv.second);
        lzp_.push_back(std::move(c));
    }
}

// mixed form: pinned members, near-singular columns (with the kernels' eigenvalue
// rounding), residual launch items
int BatchCore::plan_mixed()
{
    long so = 0, po = 0, uo = 0;
    fix_.clear();
    m00_.clear();
    m00_mem_.clear();
    ritems_.clear();
    rpitems_.clear();
    res_two_ = false;
    resp_two_ = false;
    for (size_t i = 0; i < mem.size(); i++) {
        BatchMember &m = mem[i];
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
                if (pin && zero_mode == 1 && (g.zlo_d || g.zhi_d)) pin = false;
                if (pin) {
                    m.pinned = true;
                    continue;
                }
                // (periodic z computes every column in FP64: nothing to re-solve)
                if (!g.zper && std::fabs(eig + 2.) < tau) fix_.push_back(make_int4(sx, sy, (int)i, 0));
            }
        if (m.pinned) {
            Mode00Item it{};
            it.ny = g.ny;
            it.nz = g.nz;
            it.b0 = g.zlo_d ? -2.0 - 1. : -2.0 + 1.;
            it.bn = g.zhi_d ? -2.0 - 1. : -2.0 + 1.;
            it.inv_scale = 1. / g.scale;
            // partial sums of plane 0: the residual kernel's tiles, or one per row group of
            // the fused kernel
            it.nparts = mfused_[i] ? g.ny / resx_lines(g.nx) : residual_parts(g.nx, g.ny);
            it.zper = g.zper ? 1 : 0;
            m00_.push_back(it);
            m00_mem_.push_back((int)i);
        }
        // fused members: their residual runs inside the corrections' x forward (lxc_)
        if (mfused_[i]) continue;
        // the single-solve handle's residual kernel for this member's nx (results bitwise alike)
        const bool pair = residual_pairs(g.nx);
        if (pair) resp_two_ = resp_two_ || g.ny % residual_tile_y() == 1;
        else res_two_ = res_two_ || residual_two_halo(g.nx, g.ny);
        const int zp = residual_zplanes();
        const int tx = residual_tile_x_of(g.nx), ty = residual_tile_y();
        const int gx = (g.nx + tx - 1) / tx, gy = (g.ny + ty - 1) / ty, gz = (g.nz + zp - 1) / zp;
        std::vector<int4> &ri = pair ? rpitems_ : ritems_;
        for (int bz = 0; bz < gz; bz++)
            for (int by = 0; by < gy; by++)
                for (int bx = 0; bx < gx; bx++) ri.push_back(make_int4((int)i, bx, by, bz));
    }
    nsums_ = so;
    nplane_ = po;
    nupart_ = uo;
    // the same columns as 16-column tiles, one launch class per rows-per-lane J (the PCR
    // kernel's explicit-row scheme needs the exact J of each member)
    lfix_.clear();
    {
        std::map<int, std::vector<ZItem>> cls;
        std::map<std::tuple<int, int, int>, bool> seen;
        for (const int4 &c : fix_) {
            const auto key = std::make_tuple(c.z, c.x / 16, c.y);
            if (seen.count(key)) continue;
            seen[key] = true;
            cls[(mem[c.z].g.nz + 15) / 16].push_back({c.z, c.x / 16, c.y});
        }
        for (auto &kv : cls) {
            HostClass h;
            h.key = kv.first;
            h.nitems = (int)kv.second.size();
            h.zi = std::move(kv.second);
            lfix_.push_back(std::move(h));
        }
    }
    res_.assign(mem.size(), ResArgs{});
    return PP_SUCCESS;
}

int BatchCore::plan()
{
    if (planned) return PP_CALL_ORDER;
    if (mem.empty()) return PP_INPUT_ARGS;
    long off = 0, loff = 0, doff = 0;
    lam_.clear();
    dtab_.clear();
    nz_max_ = 0;
    for (BatchMember &m : mem) {
        m.offset = off;
        off += ((long)m.g.nx * m.g.ny * m.g.nz + 63) / 64 * 64;  // 512-byte aligned fields
        nz_max_ = std::max(nz_max_, m.g.nz);
        for (int a = 0; a < 2; a++) {
            m.lam_off[a] = loff;
            const std::vector<double> l = eig_factors(m.g, a);
            lam_.insert(lam_.end(), l.begin(), l.end());
            loff += (long)l.size();
        }
        m.dtab_off = doff;
        const std::vector<double> d = dirichlet_table(m.g);
        dtab_.insert(dtab_.end(), d.begin(), d.end());
        doff += (long)d.size();
    }
    arena_elems = off;
    // FP64: z tiles (16 columns x 16 J rows, zero padding) between the y transforms and the
    // z solve, as in the single-solve handle
    ntiles_ = 0;
    if (precision == 0) {
        for (BatchMember &m : mem) {
            m.toff = -1;
            if (m.g.zper) continue;  // periodic z stays in place
            m.toff = ntiles_;
            ntiles_ += (long)((m.g.nx + 15) / 16) * 16 * m.g.ny * ((m.g.nz + 15) / 16) * 16;
        }
    }
    int rc;
    if (precision == 1) {
        if ((rc = plan_lines<float>(0, lx_)) || (rc = plan_lines<float>(1, ly_))) return rc;
        // the single-solve handle's rule for fusing a member's corrections
        mfused_.assign(mem.size(), 0);
        for (size_t i = 0; i < mem.size(); i++) {
            const SolveSetup &g = mem[i].g;
            const int lines = resx_lines(g.nx);
            mfused_[i] = fuse_res && line_compiled(g.nx) && lines > 0 && g.ny % lines == 0;
        }
        lxc_.clear();
        for (const HostClass &c : lx_)
            for (int f = 1; f >= 0; f--) {
                HostClass h;
                h.key = c.key;
                h.kind = c.kind;
                h.fused = f != 0;
                for (const LineItem &it : c.li)
                    if ((mfused_[it.solve] != 0) == h.fused) h.li.push_back(it);
                h.nitems = (int)h.li.size();
                if (h.nitems) lxc_.push_back(std::move(h));
            }
        if ((rc = plan_mixed())) return rc;
    } else {
        if ((rc = plan_lines<double>(0, lx_)) || (rc = plan_lines<double>(1, ly_))) return rc;
    }
    plan_z();
    launches[0] = (int)lx_.size();
    launches[1] = (int)ly_.size();
    launches[2] = (int)(lz_.size() + lzp_.size() + 3 * zpl_.size() + zlong_.size() + (smt_.empty() ? 0 : 1));
    ls_.assign(mem.size(), LineSolve{});
    zs_.assign(mem.size(), ZSolve{});
    // region layout
    size_t c = 0;
    o_rhs_ = take(c, (size_t)off * sizeof(double));
    o_sol_ = take(c, (size_t)off * sizeof(double));
    if (precision == 0) o_tiles_ = take(c, (size_t)ntiles_ * sizeof(double));
    if (precision == 1) o_w_ = take(c, (size_t)off * sizeof(float));
    // the solution's FP32 parts while refined (the refinement count may change after planning)
    if (precision == 1) o_u0_ = take(c, (size_t)off * sizeof(float));
    if (precision == 1) o_cacc_ = take(c, (size_t)off * sizeof(float));
    o_lam_ = take(c, lam_.size() * sizeof(double));
    o_dtab_ = take(c, dtab_.size() * sizeof(double));
    o_ls_ = take(c, ls_.size() * sizeof(LineSolve));
    o_zs_ = take(c, zs_.size() * sizeof(ZSolve));
    for (auto *v : {&lx_, &ly_, &lxc_})
        for (HostClass &h : *v) {
            h.off = take(c, h.li.size() * sizeof(LineItem));
            if (!h.li_alt.empty()) h.off_alt = take(c, h.li_alt.size() * sizeof(LineItem));
        }
    for (HostClass &h : lz_) h.off = take(c, h.zi.size() * sizeof(ZItem));
    for (HostClass &h : lzp_) h.off = take(c, h.zi.size() * sizeof(ZItem));
    o_zc2_ = take(c, zc2_.size() * sizeof(double));
    o_zth_ = take(c, zth_.size() * sizeof(ZThomasItem));
    o_zth_tab_ = take(c, zth_tab_.size() * sizeof(double));
    o_smt_ = take(c, smt_.size() * sizeof(SmtItem));
    if (precision == 1) {
        o_sums_ = take(c, (size_t)nsums_ * sizeof(double));
        o_plane_ = take(c, (size_t)nplane_ * sizeof(double));
        o_upart_ = take(c, (size_t)nupart_ * sizeof(double));
        o_fix_ = take(c, fix_.size() * sizeof(int4));
        for (HostClass &h : lfix_) h.off = take(c, h.zi.size() * sizeof(ZItem));
        o_m00_ = take(c, m00_.size() * sizeof(Mode00Item));
        o_res_ = take(c, res_.size() * sizeof(ResArgs));
        o_ritems_ = take(c, ritems_.size() * sizeof(int4));
        o_rpitems_ = take(c, rpitems_.size() * sizeof(int4));
    }
    region_bytes = (c + 255) / 256 * 256;
    planned = true;
    return PP_SUCCESS;
}

int BatchCore::bind(char *base, TableCache *tc)
{
    if (!planned || bound) return PP_CALL_ORDER;
    d_rhs = (double *)(base + o_rhs_);
    d_sol = (double *)(base + o_sol_);
    d_tiles = precision == 0 ? (double *)(base + o_tiles_) : nullptr;
    d_w = precision == 1 ? (float *)(base + o_w_) : nullptr;
    d_u0 = precision == 1 ? (float *)(base + o_u0_) : nullptr;
    d_cacc = precision == 1 ? (float *)(base + o_cacc_) : nullptr;
    d_lam_ = (double *)(base + o_lam_);
    d_dtab_ = (double *)(base + o_dtab_);
    d_ls_ = (LineSolve *)(base + o_ls_);
    d_zs_ = (ZSolve *)(base + o_zs_);
    for (auto *v : {&lx_, &ly_, &lz_, &lzp_, &lxc_})
        for (HostClass &h : *v) {
            h.d_items = base + h.off;
            h.d_items_alt = h.li_alt.empty() ? nullptr : base + h.off_alt;
        }
    d_zc2_ = (double *)(base + o_zc2_);
    d_zth_ = (ZThomasItem *)(base + o_zth_);
    d_zth_tab_ = (double *)(base + o_zth_tab_);
    d_smt_ = (SmtItem *)(base + o_smt_);
    int rc;
    for (auto *v : {&lx_, &ly_})
        for (HostClass &h : *v) {
            AxisDev t;
            if ((rc = tc->get(h.key, t))) return rc;
            tabs_[h.key] = t;
        }
    // periodic z: the complex FFT of length nz uses the twiddles of real lines of 2 nz; other
    // lengths take the real line tables of length nz
    for (HostClass &h : lzp_) {
        AxisDev t;
        if ((rc = tc->get(2 * h.key, t))) return rc;
        tabs_[2 * h.key] = t;
    }
    for (int m : zpl_) {
        AxisDev t;
        if ((rc = tc->get(mem[m].g.nz, t))) return rc;
        tabs_[mem[m].g.nz] = t;
    }
    if (precision == 1) {
        d_sums_ = (double *)(base + o_sums_);
        d_plane_ = (double *)(base + o_plane_);
        d_upart_ = (double *)(base + o_upart_);
        d_fix_ = (int4 *)(base + o_fix_);
        for (HostClass &h : lfix_) h.d_items = base + h.off;
        d_m00_ = (Mode00Item *)(base + o_m00_);
        d_res_ = (ResArgs *)(base + o_res_);
        d_ritems_ = (int4 *)(base + o_ritems_);
        d_rpitems_ = (int4 *)(base + o_rpitems_);
    }
    for (size_t i = 0; i < mem.size(); i++) {
        const BatchMember &m = mem[i];
        const SolveSetup &g = m.g;
        LineSolve &l = ls_[i];
        l.offset = m.offset;
        l.nx = g.nx;
        l.ny = g.ny;
        l.nz = g.nz;
        l.dmask = g.dmask;
        l.dz = g.dxyz[2];
        l.inv_scale = 1. / g.scale;
        l.dtab = d_dtab_ + m.dtab_off;
        l.sums_off = m.sums_off;
        l.plane_off = m.plane_off;
        l.toffset = m.toff;
        l.tnz = (precision == 0 && m.toff < 0) ? 0 : (g.nz + 15) / 16 * 16;
        ZSolve &z = zs_[i];
        z.offset = (precision == 0 && m.toff >= 0) ? m.toff : m.offset;  // FP64 solves the z tiles
        z.SX = g.nx;
        z.SY = g.ny;
        z.nz = g.nz;
        z.zlo_dirichlet = g.zlo_d;
        z.zhi_dirichlet = g.zhi_d;
        z.Cx = (g.dxyz[2] / g.dxyz[0]) * (g.dxyz[2] / g.dxyz[0]);
        z.Cy = (g.dxyz[2] / g.dxyz[1]) * (g.dxyz[2] / g.dxyz[1]);
        z.lamx = d_lam_ + m.lam_off[0];
        z.lamy = d_lam_ + m.lam_off[1];
        if (precision == 1) {
            ResArgs &r = res_[i];
            r.f = d_rhs + m.offset;
            r.u = d_sol + m.offset;
            r.u0 = d_u0 ? d_u0 + m.offset : nullptr;
            r.u1 = d_cacc ? d_cacc + m.offset : nullptr;
            r.r = d_w + m.offset;
            r.nx = g.nx;
            r.ny = g.ny;
            r.nz = g.nz;
            r.c3 = g.dxyz[2];
            r.Cx = z.Cx;
            r.Cy = z.Cy;
            for (int f = 0; f < 6; f++) r.bc[f] = g.bc[f];
            r.dtab = d_dtab_ + m.dtab_off;
            r.dmask = g.dmask;
            r.pin_defect = m.pinned ? d_plane_ + m.plane_off + g.nz : nullptr;
            r.upart = m.pinned ? d_upart_ + m.upart_off : nullptr;
        }
    }
    for (size_t k = 0; k < m00_.size(); k++) {
        const BatchMember &m = mem[m00_mem_[k]];
        m00_[k].line_sums = d_sums_ + m.sums_off;
        m00_[k].plane_add = d_plane_ + m.plane_off;
        m00_[k].upart = d_upart_ + m.upart_off;
    }
    bound = true;
    return PP_SUCCESS;
}

int BatchCore::upload(hipStream_t s)
{
    if (!bound) return PP_CALL_ORDER;
    int rc;
    // tile padding (rows from nz, columns from nx) is never written by the transforms
    if (precision == 0 && ntiles_) PPB_CHECK(hipMemsetAsync(d_tiles, 0, (size_t)ntiles_ * sizeof(double), s));
    if ((rc = put(d_lam_, lam_, s)) || (rc = put(d_dtab_, dtab_, s)) || (rc = put(d_ls_, ls_, s)) ||
        (rc = put(d_zs_, zs_, s)))
        return rc;
    for (auto *v : {&lx_, &ly_, &lxc_})
        for (HostClass &h : *v)
            if ((rc = put((LineItem *)h.d_items, h.li, s)) ||
                (!h.li_alt.empty() && (rc = put((LineItem *)h.d_items_alt, h.li_alt, s))))
                return rc;
    for (auto *v : {&lz_, &lzp_})
        for (HostClass &h : *v)
            if ((rc = put((ZItem *)h.d_items, h.zi, s))) return rc;
    if ((rc = put(d_zc2_, zc2_, s)) || (rc = put(d_smt_, smt_, s)) || (rc = put(d_zth_, zth_, s)) ||
        (rc = put(d_zth_tab_, zth_tab_, s)))
        return rc;
    if (precision == 1) {
        // members without a pinned column add zeros in the inverse-x store
        PPB_CHECK(hipMemsetAsync(d_plane_, 0, (size_t)nplane_ * sizeof(double), s));
        for (HostClass &h : lfix_)
            if ((rc = put((ZItem *)h.d_items, h.zi, s))) return rc;
        if ((rc = put(d_fix_, fix_, s)) || (rc = put(d_m00_, m00_, s)) || (rc = put(d_res_, res_, s)) ||
            (rc = put(d_ritems_, ritems_, s)) || (rc = put(d_rpitems_, rpitems_, s)))
            return rc;
    }
    return PP_SUCCESS;
}

// one stage of line transforms over every class: in/out arenas, precision form, flags
int BatchCore::run_line_stage(hipStream_t s, std::vector<HostClass> &cls, int form, int axis, int fwd, const void *in,
                              void *out, int bflags, double *sums, const double *plane, int accumulate, int ztiled,
                              const float *acc0, const float *acc1)
{
    int rc;
    for (const HostClass &c : cls)
        if ((rc = launch_class(s, c, form, axis, fwd, in, out, bflags, sums, plane, accumulate, ztiled, acc0, acc1)))
            return rc;
    return PP_SUCCESS;
}

int BatchCore::launch_class(hipStream_t s, const HostClass &c, int form, int axis, int fwd, const void *in, void *out,
                            int bflags, double *sums, const double *plane, int accumulate, int ztiled,
                            const float *acc0, const float *acc1)
{
    {
        const int v = (form == FORM_F64 && line_has_variant(c.key)) ? (force_var >= 0 ? force_var : c.var[fwd ? 0 : 1]) : 0;
        const bool alt_items = v && axis == 0;
        const AxisDev &t = tabs_.at(c.key);
        LineArgs a{};
        a.variant = v;
        a.in = in;
        a.out = out;
        a.tw = t.tw;
        a.A = t.A;
        a.B = t.B;
        a.C = t.C;
        a.D = t.D;
        a.in_scale = 1.0;
        a.out_scale = 1.0;
        a.items = (const LineItem *)(alt_items ? c.d_items_alt : c.d_items);
        a.solves = d_ls_;
        a.nlines = alt_items ? (long)c.li_alt.size() : c.nitems;
        a.bflags = bflags;
        a.line_sums = sums;
        a.plane_add = plane;
        a.accumulate = accumulate;
        a.acc0 = acc0;
        a.acc1 = acc1;
        a.ztiled = ztiled;
        // z tiles are 16 columns wide; y blocks of fewer columns fill them together
        a.tile_w = ztiled ? zsolve_cpb<double>() : 0;
        PPB_CHECK(launch_line(form, c.kind, c.key, axis, fwd, a, s));
    }
    return PP_SUCCESS;
}

int BatchCore::alt_classes() const
{
    int n = 0;
    for (const auto *v : {&lx_, &ly_})
        for (const HostClass &c : *v) n += c.var[0] + c.var[1];
    return n;
}

int BatchCore::tune_lines(hipStream_t s)
{
    if (!bound || precision != 0) return PP_CALL_ORDER;
    int rc = PP_SUCCESS;
    for (const BatchMember &m : mem)
        PPB_CHECK(launch_fill_uniform(d_rhs + m.offset, (long)m.g.nx * m.g.ny * m.g.nz, 0x7475ull, 0, s));
    if ((rc = run(s, nullptr))) return rc;  // warm-up, and every stage's input is finite
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return PP_HIP_ERROR;
    // the four FP64 line stages in run_fp64's order and arguments
    struct Stage {
        std::vector<HostClass> *cls;
        int axis, fwd;
        const void *in;
        void *out;
        int bflags, ztiled;
    };
    const Stage st[4] = {{&lx_, 0, 1, d_rhs, d_sol, BF_RHS, 0},
                         {&ly_, 1, 1, d_sol, d_tiles, 0, ZT_OUT},
                         {&ly_, 1, 0, d_tiles, d_sol, 0, ZT_IN},
                         {&lx_, 0, 0, d_sol, d_sol, BF_SCALE, 0}};
    const int saved_force = force_var;
    force_var = -1;
    for (const Stage &g : st) {
        for (HostClass &c : *g.cls) {
            if (!line_has_variant(c.key)) continue;
            float best[2] = {1e30f, 1e30f};
            for (int r = 0; r < 3 && !rc; r++) {  // shapes in alternation, best of three each
                for (int v = 0; v < 2 && !rc; v++) {
                    c.var[g.fwd ? 0 : 1] = v;
                    if (hipEventRecord(e0, s) != hipSuccess) rc = PP_HIP_ERROR;
                    if (!rc) rc = launch_class(s, c, FORM_F64, g.axis, g.fwd, g.in, g.out, g.bflags, nullptr, nullptr, 0, g.ztiled);
                    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = PP_HIP_ERROR;
                    float ms = 0.f;
                    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PP_HIP_ERROR;
                    if (!rc) best[v] = std::min(best[v], ms);
                }
            }
            c.var[g.fwd ? 0 : 1] = best[1] < 0.98f * best[0] ? 1 : 0;  // alternative only when clearly faster
        }
    }
    force_var = saved_force;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// z stage over every rows-per-lane class; FP32 with the mixed form's column rules
int BatchCore::run_z_stage(hipStream_t s, bool f32)
{
    for (const HostClass &c : lz_) {
        ZArgs z{};
        z.data = f32 ? (void *)d_w : (void *)d_tiles;
        z.tiled = f32 ? 0 : 1;
        z.zero_mode = zero_mode;
        z.items = (const ZItem *)c.d_items;
        z.solves = d_zs_;
        z.nitems = c.nitems;
        z.nz = 16 * c.key;  // selects the launch template; per-member lengths come from solves
        if (f32) {
            z.tau = tau;
            z.pinned_zero = 1;
            PPB_CHECK(launch_zsolve_pcr_f32(16 * c.key, z, s));
        } else {
            z.tau = zthomas_tau;  // the nearly singular columns are left to the Thomas sweeps
            PPB_CHECK(launch_zsolve_pcr_f64(16 * c.key, z, s));
        }
    }
    if (!f32 && !zth_.empty()) {
        ZArgs z{};
        z.data = d_tiles;
        z.tiled = 1;
        z.tile_w = zsolve_cpb<double>();
        z.tile_rows = 0;  // ceil16(nz) per member
        z.solves = d_zs_;
        PPB_CHECK(launch_zthomas_f64(z, d_zth_, d_zth_tab_, (int)zth_.size(), s));
    }
    // FP64 columns of 513..1024 rows: the long-column kernel on the member's z tiles
    for (int m : zlong_) {
        const ZSolve &zs = zs_[m];
        ZArgs z{};
        z.data = d_tiles + zs.offset;
        z.tiled = 1;
        z.tile_w = zsolve_cpb<double>();
        z.tile_rows = (zs.nz + 15) / 16 * 16;
        z.SX = zs.SX;
        z.SY = zs.SY;
        z.nz = zs.nz;
        z.lamx = d_lam_ + mem[m].lam_off[0];
        z.lamy = d_lam_ + mem[m].lam_off[1];
        z.Cx = zs.Cx;
        z.Cy = zs.Cy;
        z.zlo_dirichlet = zs.zlo_dirichlet;
        z.zhi_dirichlet = zs.zhi_dirichlet;
        z.zero_mode = zero_mode;
        PPB_CHECK(launch_zsolve_long_f64(zs.nz, z, s));
    }
    // periodic z, in place on the solution arena (FP64) or the work arena (FP32 storage, FP64
    // arithmetic, singular column zeroed)
    void *pdata = f32 ? (void *)d_w : (void *)d_sol;
    for (const HostClass &c : lzp_) {
        ZPerArgs zp{};
        zp.data = pdata;
        zp.nz = c.key;
        zp.tw = tabs_.at(2 * c.key).tw;
        zp.c2 = d_zc2_ + zc2_off_.at(c.key);
        zp.pinned_zero = f32 ? 1 : 0;
        zp.items = (const ZItem *)c.d_items;
        zp.solves = d_zs_;
        PPB_CHECK(f32 ? launch_zperiodic_f32_batch(zp, c.nitems, s) : launch_zperiodic_f64_batch(zp, c.nitems, s));
    }
    for (int m : zpl_) {
        const SolveSetup &g = mem[m].g;
        const AxisDev &t = tabs_.at(g.nz);
        const long off = mem[m].offset;
        void *data = f32 ? (void *)(d_w + off) : (void *)(d_sol + off);
        ZPerArgs zp{};
        zp.data = data;
        zp.SX = g.nx;
        zp.SY = g.ny;
        zp.nz = g.nz;
        zp.lamx = d_lam_ + mem[m].lam_off[0];
        zp.lamy = d_lam_ + mem[m].lam_off[1];
        zp.Cx = zs_[m].Cx;
        zp.Cy = zs_[m].Cy;
        zp.c2 = d_zc2_ + zc2_off_.at(g.nz);
        zp.pinned_zero = f32 ? 1 : 0;
        // z lines as the columns of a [nz][ny*nx] array: R2HC, divide, HC2R
        LineArgs a{};
        a.in = data;
        a.out = data;
        a.tw = t.tw;
        a.A = t.A;
        a.B = t.B;
        a.C = t.C;
        a.D = t.D;
        a.in_scale = 1.0;
        a.out_scale = 1.0;
        a.nlines = (long)g.nx * g.ny;
        a.ncols = g.nx * g.ny;
        a.nplanes = 1;
        const int form = f32 ? FORM_F32 : FORM_F64;
        PPB_CHECK(launch_line(form, KIND_PP, g.nz, 1, 1, a, s));
        PPB_CHECK(f32 ? launch_zper_divide_f32(zp, s) : launch_zper_divide(zp, s));
        PPB_CHECK(launch_line(form, KIND_PP, g.nz, 1, 0, a, s));
    }
    if (!f32) PPB_CHECK(launch_smt_v1_batch(d_sol, d_smt_, (int)smt_.size(), smt_nz_max_, s));
    return PP_SUCCESS;
}

int BatchCore::run_fp64(hipStream_t s, Marks *mk)
{
    int rc;
    if ((rc = run_line_stage(s, lx_, FORM_F64, 0, 1, d_rhs, d_sol, BF_RHS, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_line_stage(s, ly_, FORM_F64, 1, 1, d_sol, d_tiles, 0, nullptr, nullptr, 0, ZT_OUT))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_z_stage(s, false))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_line_stage(s, ly_, FORM_F64, 1, 0, d_tiles, d_sol, 0, nullptr, nullptr, 0, ZT_IN))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_line_stage(s, lx_, FORM_F64, 0, 0, d_sol, d_sol, BF_SCALE, nullptr, nullptr, 0))) return rc;
    return mark(s, mk);
}

// y forward, z (FP32 + FP64 near-singular columns), y inverse on the work arena
int BatchCore::mixed_inner(hipStream_t s, Marks *mk)
{
    int rc;
    if ((rc = run_line_stage(s, ly_, FORM_F32, 1, 1, d_w, d_w, 0, nullptr, nullptr, 0))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_z_stage(s, true))) return rc;
    if (!fix_.empty()) {
        ZArgs z{};
        z.data = d_w;
        z.solves = d_zs_;
        z.zero_mode = zero_mode;
        z.tau = tau;
        z.pinned_zero = 1;
        if (zfix_tiles_supported(nz_max_)) {
            for (const HostClass &h : lfix_) {
                z.items = (const ZItem *)h.d_items;
                z.nitems = h.nitems;
                PPB_CHECK(launch_zfix_tiles_f32(16 * h.key, z, s));
            }
        } else {
            PPB_CHECK(launch_zfix64_f32(z, d_fix_, (int)fix_.size(), nz_max_, s));
        }
    }
    if ((rc = mark(s, mk))) return rc;
    if ((rc = run_line_stage(s, ly_, FORM_F32, 1, 0, d_w, d_w, 0, nullptr, nullptr, 0))) return rc;
    return mark(s, mk);
}

// the mixed form over the batch (same steps and solution parts as the single-solve handle's)
int BatchCore::run_mixed(hipStream_t s, Marks *mk)
{
    int rc;
    const int nm00 = (int)m00_.size();
    double *sums = nm00 ? d_sums_ : nullptr;
    const double *plane = nm00 ? d_plane_ : nullptr;
    const int R = refine;
    if ((rc = run_line_stage(s, lx_, FORM_F64_TO_F32, 0, 1, d_rhs, d_w, BF_RHS, sums, nullptr, 0))) return rc;
    if ((rc = mark(s, mk))) return rc;
    if ((rc = mixed_inner(s, mk))) return rc;
    if (nm00) PPB_CHECK(launch_mode00_batch(d_m00_, nm00, nz_max_, 0, s));
    // u0 in FP32 while refinement follows, else the FP64 solution
    if (R > 0) {
        if ((rc = run_line_stage(s, lx_, FORM_F32, 0, 0, d_w, d_u0, BF_SCALE, nullptr, plane, 0))) return rc;
    } else {
        if ((rc = run_line_stage(s, lx_, FORM_F32_TO_F64, 0, 0, d_w, d_sol, BF_SCALE, nullptr, plane, 0))) return rc;
    }
    if ((rc = mark(s, mk))) return rc;
    for (int it = 0; it < R; it++) {
        const int src = it == 0 ? RES_U0 : RES_U0_U1;
        PPB_CHECK(launch_residual_f32_batch(d_res_, d_ritems_, (int)ritems_.size(), res_two_, src, s));
        PPB_CHECK(launch_residual_pair_f32_batch(d_res_, d_rpitems_, (int)rpitems_.size(), resp_two_, src, s));
        if ((rc = mark(s, mk))) return rc;
        // the corrections' x forward: fused members compute their residual inside it
        for (const HostClass &c : lxc_) {
            if (c.fused) {
                const AxisDev &t = tabs_.at(c.key);
                LineArgs a{};
                a.in = d_w;
                a.out = d_w;
                a.tw = t.tw;
                a.A = t.A;
                a.B = t.B;
                a.C = t.C;
                a.D = t.D;
                a.in_scale = 1.0;
                a.out_scale = 1.0;
                a.items = (const LineItem *)c.d_items;
                a.solves = d_ls_;
                a.nlines = c.nitems;
                a.line_sums = sums;
                PPB_CHECK(launch_resx(c.kind, c.key, src, a, ResArgs{}, d_res_, s));
            } else if ((rc = launch_class(s, c, FORM_F32, 0, 1, d_w, d_w, 0, sums, nullptr, 0, 0))) {
                return rc;
            }
        }
        if ((rc = mark(s, mk))) return rc;
        if ((rc = mixed_inner(s, mk))) return rc;
        if (nm00) PPB_CHECK(launch_mode00_batch(d_m00_, nm00, nz_max_, 1, s));
        if (it == R - 1) {
            // u = u0 + csum + c, in FP64
            if ((rc = run_line_stage(s, lx_, FORM_F32_TO_F64, 0, 0, d_w, d_sol, BF_SCALE, nullptr, plane, 0, 0, d_u0,
                                     it > 0 ? d_cacc : nullptr)))
                return rc;
        } else {
            // csum = c1, then csum += c
            if ((rc = run_line_stage(s, lx_, FORM_F32, 0, 0, d_w, d_cacc, BF_SCALE, nullptr, plane, it > 0 ? 1 : 0)))
                return rc;
        }
        if ((rc = mark(s, mk))) return rc;
    }
    return PP_SUCCESS;
}

int BatchCore::run(hipStream_t s, Marks *mk)
{
    if (!bound) return PP_CALL_ORDER;
    int rc = mark(s, mk);
    if (rc) return rc;
    return precision == 1 ? run_mixed(s, mk) : run_fp64(s, mk);
}

}  // namespace pp

using namespace pp;

struct pp_batch_s {
    BatchCore core;
    TableCache tables;
    char *region = nullptr;
    bool finalized = false;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // timing: events between the launches of the last solves
    bool timing = false;
    static constexpr int kMarks = 64;
    std::vector<hipEvent_t> ev;
    int ev_slots = 0, ev_next = 0, ev_used = 0, marks_last = 0;
    // PP_OPT_GRAPH: the solve's launches captured once into a HIP graph and replayed (the
    // first solve runs eagerly, so one-time launch set-up happens outside the capture)
    int graph_mode = 1;
    int autotune = 0;  // PP_OPT_AUTOTUNE: FP64 line classes tuned at finalize
    int solves = 0;
    hipGraphExec_t graph = nullptr;
    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
};

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
    b->drop_graph();
    if (b->region) (void)hipFree(b->region);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return PP_SUCCESS;
}

int pp_batch_add(pp_batch b, int nx, int ny, int nz, const char *bc, const double *X, int *id)
{
    if (!b) return PP_INVALID_HANDLE;
    if (b->finalized) return PP_CALL_ORDER;
    return b->core.add(nx, ny, nz, bc, X, id);
}

int pp_batch_finalize(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    if (b->finalized) return PP_CALL_ORDER;
    int rc = b->core.plan();
    if (rc) return rc;
    if (hipMalloc((void **)&b->region, b->core.region_bytes) != hipSuccess) {
        b->region = nullptr;
        return PP_ALLOCATION_FAIL;
    }
    if ((rc = b->core.bind(b->region, &b->tables)) || (rc = b->core.upload(b->stream))) return rc;
    PPB_CHECK(hipStreamSynchronize(b->stream));
    if (b->autotune && b->core.precision == 0) {
        if ((rc = b->core.tune_lines(b->stream))) return rc;
        PPB_CHECK(hipStreamSynchronize(b->stream));
    }
    b->finalized = true;
    return PP_SUCCESS;
}

int pp_batch_tuning(pp_batch b, int *alt_shapes)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (!alt_shapes) return PP_INPUT_ARGS;
    *alt_shapes = b->core.alt_classes();
    return PP_SUCCESS;
}

int pp_batch_count(pp_batch b, int *members, long long *points)
{
    if (!b) return PP_INVALID_HANDLE;
    if (members) *members = (int)b->core.mem.size();
    if (points) *points = b->core.points();
    return PP_SUCCESS;
}

int pp_batch_launch_count(pp_batch b, int *per_stage)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (!per_stage) return PP_INPUT_ARGS;
    for (int i = 0; i < 3; i++) per_stage[i] = b->core.launches[i];
    return PP_SUCCESS;
}

int pp_batch_field(pp_batch b, int id, double **d_rhs)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (id < 0 || id >= (int)b->core.mem.size() || !d_rhs) return PP_INPUT_ARGS;
    *d_rhs = b->core.d_rhs + b->core.mem[id].offset;
    return PP_SUCCESS;
}

int pp_batch_solution(pp_batch b, int id, double **d_u)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    if (id < 0 || id >= (int)b->core.mem.size() || !d_u) return PP_INPUT_ARGS;
    *d_u = b->core.d_sol + b->core.mem[id].offset;
    return PP_SUCCESS;
}

static size_t member_bytes(pp_batch b, int id)
{
    const SolveSetup &g = b->core.mem[id].g;
    return (size_t)g.nx * g.ny * g.nz * sizeof(double);
}

int pp_batch_set_rhs(pp_batch b, int id, const double *rhs)
{
    double *d = nullptr;
    int rc = pp_batch_field(b, id, &d);
    if (rc) return rc;
    if (!rhs) return PP_INPUT_ARGS;
    PPB_CHECK(hipMemcpyAsync(d, rhs, member_bytes(b, id), hipMemcpyHostToDevice, b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_get_solution(pp_batch b, int id, double *u)
{
    double *d = nullptr;
    int rc = pp_batch_solution(b, id, &d);
    if (rc) return rc;
    if (!u) return PP_INPUT_ARGS;
    PPB_CHECK(hipMemcpyAsync(u, d, member_bytes(b, id), hipMemcpyDeviceToHost, b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_get_rhs(pp_batch b, int id, double *rhs)
{
    double *d = nullptr;
    int rc = pp_batch_field(b, id, &d);
    if (rc) return rc;
    if (!rhs) return PP_INPUT_ARGS;
    PPB_CHECK(hipMemcpyAsync(rhs, d, member_bytes(b, id), hipMemcpyDeviceToHost, b->stream));
    PPB_CHECK(hipStreamSynchronize(b->stream));
    return PP_SUCCESS;
}

int pp_batch_fill_uniform(pp_batch b, unsigned long long seed, int round_f32)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    for (size_t i = 0; i < b->core.mem.size(); i++) {
        const BatchMember &m = b->core.mem[i];
        PPB_CHECK(launch_fill_uniform(b->core.d_rhs + m.offset, (long)m.g.nx * m.g.ny * m.g.nz, seed + i, round_f32,
                                      b->stream));
    }
    return PP_SUCCESS;
}

int pp_batch_set_stream(pp_batch b, void *stream)
{
    if (!b) return PP_INVALID_HANDLE;
    PPB_CHECK(hipStreamSynchronize(b->stream));
    b->drop_graph();
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
        b->core.zero_mode = (int)value;
        return PP_SUCCESS;
    case PP_OPT_PRECISION:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        if (b->finalized || !b->core.mem.empty()) return PP_CALL_ORDER;
        b->core.precision = (int)value;
        return PP_SUCCESS;
    case PP_OPT_REFINE:
        if (value < 0 || value > 16) return PP_INPUT_ARGS;
        if (b->core.refine != (int)value) {
            PPB_CHECK(hipStreamSynchronize(b->stream));
            b->drop_graph();
        }
        b->core.refine = (int)value;
        return PP_SUCCESS;
    case PP_OPT_AUTOTUNE:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        if (b->finalized) return PP_CALL_ORDER;
        b->autotune = (int)value;
        return PP_SUCCESS;
    case PP_OPT_FUSE_RESIDUAL:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        if (b->finalized) return PP_CALL_ORDER;
        b->core.fuse_res = (int)value;
        return PP_SUCCESS;
    case PP_OPT_LINE_VARIANT:
        if (value < -1 || value > 1) return PP_INPUT_ARGS;
        PPB_CHECK(hipStreamSynchronize(b->stream));
        b->drop_graph();
        b->core.force_var = (int)value;
        return PP_SUCCESS;
    case PP_OPT_GRAPH:
        if (value != 0 && value != 1) return PP_INPUT_ARGS;
        PPB_CHECK(hipStreamSynchronize(b->stream));
        b->drop_graph();
        b->graph_mode = (int)value;
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
    if ((key != PP_OPT_TAU && key != PP_OPT_ZTHOMAS_TAU) || !(value >= 0.0 && value < 4.0)) return PP_INPUT_ARGS;
    if (b->finalized) return PP_CALL_ORDER;
    if (key == PP_OPT_TAU) b->core.tau = value;
    else b->core.zthomas_tau = value;
    return PP_SUCCESS;
}

// solve every member: right-hand sides kept, solutions in the solution arena; each stage is
// one launch per shape class
int pp_batch_solve(pp_batch b)
{
    if (!b) return PP_INVALID_HANDLE;
    if (!b->finalized) return PP_CALL_ORDER;
    Marks mk;
    if (b->timing) {
        mk.ev = &b->ev[(size_t)pp_batch_s::kMarks * b->ev_next];
        mk.max = pp_batch_s::kMarks;
        b->ev_next = (b->ev_next + 1) % b->ev_slots;
        if (b->ev_used < b->ev_slots) b->ev_used++;
    }
    const bool use_graph = b->graph_mode && !b->timing && b->stream != nullptr;
    if (use_graph && b->graph) {
        PPB_CHECK(hipGraphLaunch(b->graph, b->stream));
        b->solves++;
        return PP_SUCCESS;
    }
    if (use_graph && b->solves > 0) {
        // capture this solve's launches (kernels only: no memset or copy is recorded) and
        // replay the graph from now on
        hipGraph_t g = nullptr;
        PPB_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        const int rc = b->core.run(b->stream, nullptr);
        const hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (rc || ec != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : PP_HIP_ERROR;
        }
        const hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            b->graph = nullptr;
            return PP_HIP_ERROR;
        }
        PPB_CHECK(hipGraphLaunch(b->graph, b->stream));
        b->solves++;
        return PP_SUCCESS;
    }
    const int rc = b->core.run(b->stream, &mk);
    if (mk.ev) b->marks_last = mk.n;
    if (!rc) b->solves++;
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
