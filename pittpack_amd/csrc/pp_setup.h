// pp_setup.h -- host-side set-up of one solve, shared by the single-solve handle and the
// batched form: boundary conditions -> transform kinds, Dirichlet tags, eigenvalue
// parameters and scale (assignBoundary / detectTransforms / extractTag / setEigenVal /
// setScale, src/pencilDcmp.cpp:3533-3732), grid spacing (setDxyz, 826-840), the eigenvalue
// factors (getEigenVal, 4393-4417), the Dirichlet boundary-term tables (modifyRhsDirichlet
// + exactValue, 4623-4835 and 1837-1866) and the FP64 twiddle tables of a line length.
#pragma once
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "pittpack_hip.h"
#include "pp_line.h"
#include "pp_line_bs.h"

namespace pp {

constexpr double kPi = 3.1415926535897932384;  // src/include/params.h:116

// exactValue (src/pencilDcmp.cpp:1837-1866)
inline double exact_value(double omega, double x, int tg)
{
    double ans = 0.0;
    if (std::fabs(omega) < 1.0e-6) return 1.0;
    if (tg == 0) ans = std::sin(omega * x);
    else if (tg == 1) ans = std::cos(omega * x);
    else if (tg == -1) ans = std::sin(omega * x);
    else if (tg == -2) ans = std::sin(omega * x);
    return ans;
}

inline int axis_kind(char lo, char hi)
{
    if (lo == 'D' && hi == 'D') return KIND_DD;
    if (lo == 'N' && hi == 'N') return KIND_NN;
    if (lo == 'P' && hi == 'P') return KIND_PP;
    if (lo == 'N' && hi == 'D') return KIND_ND;
    if (lo == 'D' && hi == 'N') return KIND_DN;
    return -1;
}

// per-axis twiddle tables in FP64 (host) for a real line of length N
struct AxisTables {
    std::vector<double> tw, A, B, C, D;  // interleaved complex
};

// sizes (doubles) of the five tables of line length N: tw, A, B, C, D
struct AxisTableSizes {
    size_t n[5];
};

inline AxisTableSizes axis_table_sizes(int N)
{
    if (!line_compiled(N)) {
        const size_t L = (size_t)bs_len(N);
        return {{2 * L, 12 * (size_t)N, 8 * L, 12 * (size_t)N, 2}};
    }
    const size_t M = (size_t)N / 2;
    return {{2 * M, 2 * (M + 1), 2 * (M + 1), 2 * M, 2 * M}};
}

// Tables of the arbitrary-length transforms (pp_line_bs.h) for line length N, convolution
// length L = bs_len(N), written to caller arrays of axis_table_sizes(N):
//   tw = exp(-2 pi i e / L), e < L (Stockham twiddles);
//   A  = P tables, C = Q tables: family {DCT/DST-II/III, DCT/DST-IV, real DFT} x {forward,
//        inverse}, N complex each (pre- and post-twiddles with the chirp folded in);
//   B  = FFT_L of the chirp kernel / L: chirp {half-angle (beta = 1/N), full (beta = 2/N)} x
//        {forward, inverse}, L complex each.
// Angles of chirp values are reduced exactly in integers (m^2 mod 4N or 2N) and evaluated in
// long double; the kernel spectra come from a long-double radix-2 FFT.
inline void fill_bs_tables(int N, double *tw, double *A, double *B, double *C, double *D)
{
    using ld = long double;
    const ld pi = 3.141592653589793238462643383279502884L;
    const int L = bs_len(N);
    auto cis = [](ld a) { return std::pair<ld, ld>(std::cos(a), std::sin(a)); };
    // exp(-2 pi i k / L), k < L / 2: the FFT twiddles of every section and, with
    // exp(-2 pi i (k + L/2) / L) = -exp(-2 pi i k / L), the Stockham table
    static thread_local ld wr[kBsMaxN], wi[kBsMaxN];
    for (int k = 0; k < L / 2; k++) {
        const auto w = cis(-2 * pi * (ld)k / (ld)L);
        wr[k] = w.first;
        wi[k] = w.second;
    }
    for (int e = 0; e < L; e++) {
        const ld sg = e < L / 2 ? 1 : -1;
        tw[2 * e] = (double)(sg * wr[e % (L / 2)]);
        tw[2 * e + 1] = (double)(sg * wi[e % (L / 2)]);
    }
    // chirp c_m = exp(i pi beta m^2 / 2): half-angle exp(i pi m^2 / (2N)), full exp(i pi m^2 / N),
    // m < N, each evaluated once
    static thread_local ld chr[2][kBsMaxN], chi[2][kBsMaxN];
    for (int m = 0; m < N; m++)
        for (int full = 0; full < 2; full++) {
            const long long per = full ? 2LL * N : 4LL * N;
            const long long r = ((long long)m * m) % per;
            const auto c = cis(full ? pi * (ld)r / (ld)N : pi * (ld)r / (2 * (ld)N));
            chr[full][m] = c.first;
            chi[full][m] = c.second;
        }
    auto chirp = [&](int m, bool full) { return std::pair<ld, ld>(chr[full][m], chi[full][m]); };
    // kernel spectra: b_j = c_j (j < N), b_{L-j} = c_j (0 < j < N), zero elsewhere; inverse
    // direction uses conj(c)
    static thread_local ld re[2 * kBsMaxN], im[2 * kBsMaxN];
    for (int sec = 0; sec < 4; sec++) {
        const bool full = sec >= 2, inv = sec & 1;
        for (int j = 0; j < L; j++) re[j] = im[j] = 0;
        for (int j = 0; j < N; j++) {
            const auto c = chirp(j, full);
            const ld ci = inv ? -c.second : c.second;
            re[j] = c.first;
            im[j] = ci;
            if (j > 0) {
                re[L - j] = c.first;
                im[L - j] = ci;
            }
        }
        // in-place radix-2 FFT (forward, exp(-2 pi i jk / L)), long double
        for (int i = 1, j = 0; i < L; i++) {
            int bit = L >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                std::swap(re[i], re[j]);
                std::swap(im[i], im[j]);
            }
        }
        for (int len = 2; len <= L; len <<= 1)
            for (int i = 0; i < L; i += len)
                for (int k = 0; k < len / 2; k++) {
                    // exp(-2 pi i k / len) = twiddle k L / len of the length-L table
                    const ld cr = wr[k * (L / len)], ci = wi[k * (L / len)];
                    const ld xr = re[i + k + len / 2] * cr - im[i + k + len / 2] * ci;
                    const ld xi = re[i + k + len / 2] * ci + im[i + k + len / 2] * cr;
                    re[i + k + len / 2] = re[i + k] - xr;
                    im[i + k + len / 2] = im[i + k] - xi;
                    re[i + k] += xr;
                    im[i + k] += xi;
                }
        for (int j = 0; j < L; j++) {
            B[2 * ((size_t)sec * L + j)] = (double)(re[j] / (ld)L);
            B[2 * ((size_t)sec * L + j) + 1] = (double)(im[j] / (ld)L);
        }
    }
    // pre (P) and post (Q) twiddles; forward: a_j = x_j pre_j conj(c_j), out from
    // post_k conj(c_k) S'_k; inverse: pre_j c_j and post_k c_k
    auto put = [&](double *v, int sec, int j, std::pair<ld, ld> a, std::pair<ld, ld> b) {
        v[2 * ((size_t)sec * N + j)] = (double)(a.first * b.first - a.second * b.second);
        v[2 * ((size_t)sec * N + j) + 1] = (double)(a.first * b.second + a.second * b.first);
    };
    for (int j = 0; j < N; j++) {
        const auto ch = chirp(j, false), cf = chirp(j, true);
        const std::pair<ld, ld> chc = {ch.first, -ch.second}, cfc = {cf.first, -cf.second};
        const std::pair<ld, ld> one = {1, 0};
        // DCT/DST-II forward: a = x, out = Re(2 exp(-i pi k/(2N)) S_k)
        const auto h = cis(-pi * (ld)j / (2 * (ld)N));
        put(A, 0, j, one, chc);
        put(C, 0, j, {2 * h.first, 2 * h.second}, chc);
        // DCT/DST-III/2 inverse: a = w_j exp(+i pi j/(2N)) X_j (w_0 = 1/2), out = Re(S_n)
        const ld wj = j == 0 ? 0.5L : 1.0L;
        put(A, 1, j, {wj * h.first, -wj * h.second}, ch);
        put(C, 1, j, one, ch);
        // DCT/DST-IV: a = exp(-i pi j/(2N)) x_j, out = f exp(-i pi (2k+1)/(4N)) S_k, f = 2 / 1
        const auto q = cis(-pi * (2 * (ld)j + 1) / (4 * (ld)N));
        put(A, 2, j, h, chc);
        put(C, 2, j, {2 * q.first, 2 * q.second}, chc);
        put(A, 3, j, h, chc);
        put(C, 3, j, q, chc);
        // real DFT: forward exp(-2 pi i jk/N), inverse exp(+2 pi i jk/N), unnormalised
        put(A, 4, j, one, cfc);
        put(C, 4, j, one, cfc);
        put(A, 5, j, one, cf);
        put(C, 5, j, one, cf);
    }
    D[0] = D[1] = 0.0;
}

// the five tables of line length N into caller arrays of axis_table_sizes(N)
inline void fill_axis_tables(int N, double *tw, double *A, double *B, double *C, double *D)
{
    if (!line_compiled(N)) {
        fill_bs_tables(N, tw, A, B, C, D);
        return;
    }
    const int M = N / 2;
    auto put = [](double *v, int cnt, auto f) {
        for (int k = 0; k < cnt; k++) {
            double ang = f(k);
            v[2 * k] = std::cos(ang);
            v[2 * k + 1] = std::sin(ang);
        }
    };
    put(tw, M, [&](int e) { return -2.0 * kPi * (double)e / (double)M; });
    put(A, M + 1, [&](int k) { return -2.0 * kPi * (double)k / (double)N; });
    put(B, M + 1, [&](int k) { return -kPi * (double)k / (2.0 * N); });
    put(C, M, [&](int n) { return -kPi * (4.0 * n + 1.0) / (4.0 * N); });
    put(D, M, [&](int k) { return -kPi * (double)k / (double)N; });
}

inline void make_axis_tables(int N, AxisTables &t)
{
    const AxisTableSizes z = axis_table_sizes(N);
    std::vector<double> *v[5] = {&t.tw, &t.A, &t.B, &t.C, &t.D};
    for (int i = 0; i < 5; i++) v[i]->assign(z.n[i], 0.0);
    fill_axis_tables(N, t.tw.data(), t.A.data(), t.B.data(), t.C.data(), t.D.data());
}

// the host tables of line length N, built once per process and shared by every handle, batch
// and server (an arbitrary-length table set takes about a millisecond to build)
inline const AxisTables &axis_tables_cached(int N)
{
    static std::mutex mu;
    static std::map<int, std::unique_ptr<AxisTables>> cache;
    std::lock_guard<std::mutex> lk(mu);
    std::unique_ptr<AxisTables> &p = cache[N];
    if (!p) {
        p.reset(new AxisTables);
        make_axis_tables(N, *p);
    }
    return *p;
}

// everything one solve's kernels need from (sizes, bc, box)
struct SolveSetup {
    int nx = 0, ny = 0, nz = 0;
    char bc[6] = {0, 0, 0, 0, 0, 0};
    double X[6] = {0, 1, 0, 1, 0, 1};
    int kind[2] = {-1, -1};
    int zlo_d = 0, zhi_d = 0;
    int zper = 0;  // periodic z (both z faces P)
    int dtag[3] = {0, 0, 0};
    double num[2] = {0, 0}, denum[2] = {0, 0}, freqs[2] = {kPi, kPi};
    double scale = 1.0;
    double dxyz[3] = {0, 0, 0};
    int dmask = 0;
};

// assignBoundary + detectTransforms + extractTag + setEigenVal + setScale
inline int setup_boundary(SolveSetup &g, const char *bc)
{
    for (int i = 0; i < 6; i++)
        if (bc[i] != 'P' && bc[i] != 'N' && bc[i] != 'D') return PP_UNSUPPORTED_BC;
    const int kx = axis_kind(bc[0], bc[1]), ky = axis_kind(bc[2], bc[3]);
    if (kx < 0 || ky < 0) return PP_UNSUPPORTED_BC;
    if ((bc[4] == 'P') != (bc[5] == 'P')) return PP_UNSUPPORTED_BC;  // one periodic z face
    for (int i = 0; i < 6; i++) g.bc[i] = bc[i];
    g.zper = bc[4] == 'P';
    g.kind[0] = kx;
    g.kind[1] = ky;
    g.zlo_d = bc[4] == 'D';
    g.zhi_d = bc[5] == 'D';
    // detectTransforms + extractTag (src/pencilDcmp.cpp:3636-3732): tags used by the
    // Dirichlet boundary term (exactValue's tg)
    char tr[3];
    for (int i = 0; i < 3; i++) {
        const char lo = bc[2 * i], hi = bc[2 * i + 1];
        tr[i] = 'U';
        if (lo == 'D' && hi == 'D') tr[i] = 'S';
        else if (lo == 'N' && hi == 'N') tr[i] = 'C';
        else if (lo == 'P' && hi == 'P') tr[i] = 'F';
    }
    int count = 0;
    for (int i = 0; i < 3; i++) count += (tr[i] == 'U');
    if (count == 0) tr[2] = 'U';
    for (int i = 0; i < 3; i++) g.dtag[i] = tr[i] == 'S' ? 0 : tr[i] == 'C' ? 1 : tr[i] == 'F' ? 2 : -1;
    if (bc[4] == 'D') g.dtag[2] = -1;
    else if (bc[4] == 'N') g.dtag[2] = -2;
    // setEigenVal (3533-3576) with num = 1/2 for the quarter-wave axes
    const int k2[2] = {kx, ky};
    for (int a = 0; a < 2; a++) {
        g.freqs[a] = kPi;
        g.denum[a] = 0.0;
        switch (k2[a]) {
        case KIND_DD: g.num[a] = 1.0; break;
        case KIND_NN: g.num[a] = 0.0; break;
        case KIND_PP: g.num[a] = 0.0; g.freqs[a] = 2. * g.freqs[a]; break;
        default: g.num[a] = 0.5; break;
        }
    }
    // setScale (3670-3683)
    g.scale = (double)g.nx * (double)g.ny;
    g.dmask = 0;
    for (int f = 0; f < 6; f++)
        if (bc[f] == 'D') g.dmask |= (1 << f);
    return PP_SUCCESS;
}

// setDxyz (src/pencilDcmp.cpp:826-840); at one rank every orientation spans the whole box
inline void setup_coords(SolveSetup &g)
{
    const double *X = g.X;
    g.dxyz[0] = -(X[0] - X[1]) / (g.nx);
    g.dxyz[1] = -(X[2] - X[3]) / (g.ny);
    g.dxyz[2] = -(X[4] - X[5]) / (g.nz);
}

// eigenvalue factors -2 + 2 cos((j + num) * freq / (n + denum)) (src/pencilDcmp.cpp:4410-4413)
inline void fill_eig_factors(const SolveSetup &g, int axis, double *l)
{
    const int n = axis == 0 ? g.nx : g.ny;
    for (int j = 0; j < n; j++) l[j] = (-2.0 + 2. * std::cos((j + g.num[axis]) * g.freqs[axis] / (n + g.denum[axis])));
}

inline std::vector<double> eig_factors(const SolveSetup &g, int axis)
{
    std::vector<double> l(axis == 0 ? g.nx : g.ny);
    fill_eig_factors(g, axis, l.data());
    return l;
}

// Dirichlet term tables (modifyRhsDirichlet, src/pencilDcmp.cpp:4623-4835, SHIFT = 1):
// [ex(nx) ey(ny) ez(nz) x-faces(2) y-faces(2) z-faces(2)]
inline size_t dirichlet_table_size(const SolveSetup &g) { return (size_t)g.nx + g.ny + g.nz + 6; }

inline void fill_dirichlet_table(const SolveSetup &g, double *dt)
{
    const double pi = 4. * std::atan(1.0);
    const double omega[3] = {1 * pi, 2 * pi, 3 * pi};
    const double Xa = g.X[0], Ya = g.X[2], Za = g.X[4];
    const double c1 = g.dxyz[0], c2 = g.dxyz[1], c3 = g.dxyz[2];
    const double shift = 1.0;
    size_t o = 0;
    for (int i = 0; i < g.nx; i++) dt[o++] = exact_value(omega[0], Xa + i * c1 + shift * c1 * .5, g.dtag[0]);
    for (int j = 0; j < g.ny; j++) dt[o++] = exact_value(omega[1], Ya + j * c2 + shift * c2 * 0.5, g.dtag[1]);
    for (int k = 0; k < g.nz; k++) dt[o++] = exact_value(omega[2], Za + k * c3 + shift * c3 * .5, g.dtag[2]);
    dt[o++] = exact_value(omega[0], Xa, g.dtag[0]);
    dt[o++] = exact_value(omega[0], Xa + ((g.nx - 1) + 1) * c1, g.dtag[0]);
    dt[o++] = exact_value(omega[1], Ya, g.dtag[1]);
    dt[o++] = exact_value(omega[1], Ya + ((g.ny - 1) + 1) * c2, g.dtag[1]);
    dt[o++] = exact_value(omega[2], Za, g.dtag[2]);
    dt[o++] = exact_value(omega[2], Za + (g.nz - 1) * c3 + c3, g.dtag[2]);
}

inline std::vector<double> dirichlet_table(const SolveSetup &g)
{
    std::vector<double> dt(dirichlet_table_size(g));
    fill_dirichlet_table(g, dt.data());
    return dt;
}

// periodic z: 2 cos(2 pi k / nz), k < nz (eigenvalues of the circulant z operator minus eig)
inline void fill_zper_cos2(int nz, double *c)
{
    for (int k = 0; k < nz; k++) c[k] = 2.0 * std::cos(2.0 * kPi * (double)k / (double)nz);
}

inline std::vector<double> zper_cos2(int nz)
{
    std::vector<double> c(nz);
    fill_zper_cos2(nz, c.data());
    return c;
}

// eigenvalue of column (sx, sy) with the kernels' rounding (no contraction)
inline double column_eig(const std::vector<double> &lx, const std::vector<double> &ly, int sx, int sy, double Cx,
                         double Cy)
{
    volatile double t1 = lx[sx] * Cx;
    volatile double t2 = -2.0 + t1;
    volatile double t3 = ly[sy] * Cy;
    return t2 + t3;
}

}  // namespace pp
