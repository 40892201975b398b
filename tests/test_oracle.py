"""Pins the CPU oracle (oracle/pp_oracle.c) before it is trusted as the GPU checker.

The reference ships no tests, golden vectors or fixtures and cannot be built here, so each
oracle function is checked against something independent of it:
  * FFT and the line transforms: long-double brute-force sums from the definitions
    (tests/golden/transforms.npz) and scipy.fft (pocketfft);
  * z solves: dense LAPACK solves of the same tridiagonal / cyclic systems;
  * the whole solve: a dense 7-point ghost-cell operator assembled here with Kronecker
    sums, the discrete residual, the (0,0)-mode row-0 rule in its physical form, and
    second-order convergence on manufactured solutions.
A planted mistake (sign, index, dropped term) in any oracle step fails at least one test.
"""
import os

import numpy as np
import pytest
import scipy.fft as sf

import oracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "transforms.npz"))
SIZES = [8, 12, 16, 17, 24, 48, 64, 96]


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("n", SIZES)
def test_fft_known_answers(n):
    x = GOLD[f"x_{n}"]
    z = oracle.fft(x.astype(complex), -1)
    assert rel(z.real, GOLD[f"dft_re_{n}"]) < 1e-14
    assert rel(z.imag, GOLD[f"dft_im_{n}"]) < 1e-13 or np.abs(GOLD[f"dft_im_{n}"]).max() < 1e-12
    # backward = conjugate transform, unnormalised
    w = oracle.fft(z, +1)
    assert rel(w.real, n * x) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 5, 7, 30, 37, 101, 128, 210])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(oracle.fft(x, -1), np.fft.fft(x)) < 1e-13
    assert rel(oracle.fft(x, +1), np.fft.ifft(x) * n) < 1e-13


@pytest.mark.parametrize("n", SIZES)
def test_reference_dct_dst_known_answers(n):
    """signalProc's DCT10 / DST10 (Makhoul + FFT) equal REDFT10 / RODFT10 by definition."""
    x = GOLD[f"x_{n}"]
    assert rel(oracle.line_transform(x, oracle.DCT2), GOLD[f"dct2_{n}"]) < 1e-14
    assert rel(oracle.line_transform(x, oracle.DST2), GOLD[f"dst2_{n}"]) < 1e-14
    assert rel(oracle.line_transform(x, oracle.DCT4), GOLD[f"dct4_{n}"]) < 1e-14
    assert rel(oracle.line_transform(x, oracle.DST4), GOLD[f"dst4_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [8, 9, 16, 33, 64, 100])
def test_line_transforms_vs_scipy(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    assert rel(oracle.line_transform(x, oracle.DCT2), sf.dct(x, 2)) < 1e-13
    assert rel(oracle.line_transform(x, oracle.DST2), sf.dst(x, 2)) < 1e-13
    # inverse steps: reference DCT01 / DST01 are half of DCT-III / DST-III
    assert rel(oracle.line_transform(x, oracle.DCT3H), sf.dct(x, 3) / 2) < 1e-13
    assert rel(oracle.line_transform(x, oracle.DST3H), sf.dst(x, 3) / 2) < 1e-13
    assert rel(oracle.line_transform(x, oracle.DCT4), sf.dct(x, 4)) < 1e-13
    assert rel(oracle.line_transform(x, oracle.DST4), sf.dst(x, 4)) < 1e-13
    # round trips are N (reference scale = Nx*Ny, src/pencilDcmp.cpp:3670-3683)
    for fwd, inv in ((oracle.DCT2, oracle.DCT3H), (oracle.DST2, oracle.DST3H),
                     (oracle.DCT4, oracle.DCT4H), (oracle.DST4, oracle.DST4H)):
        y = oracle.line_transform(oracle.line_transform(x, fwd), inv)
        assert rel(y, n * x) < 1e-13


def tri_dense(n, eig, bcz, pin):
    A = np.diag(np.full(n, eig)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    A[0, 0] = eig - 1 if bcz[0] == "D" else eig + 1
    A[-1, -1] = eig - 1 if bcz[1] == "D" else eig + 1
    if pin:
        A[0, :] = 0
        A[0, 0] = 1
    return A


@pytest.mark.parametrize("bcz", ["DD", "NN", "DN", "ND"])
@pytest.mark.parametrize("eig", [-2.0, -2.0 - 3e-4, -2.7, -9.5])
@pytest.mark.parametrize("n", [5, 16, 63])
def test_zsolve_line_vs_dense(n, eig, bcz):
    rng = np.random.default_rng(n)
    r = rng.standard_normal(n)
    for zero_mode in (0, 1):
        pin = abs(eig + 2) < 1e-10 and (zero_mode == 0 or bcz == "NN")
        if abs(eig + 2) < 1e-10 and not pin and bcz == "NN":
            continue
        ref = np.linalg.solve(tri_dense(n, eig, bcz, pin), r)
        got = oracle.zsolve_line(r, eig, bcz, zero_mode)
        assert rel(got, ref) < 1e-11, (n, eig, bcz, zero_mode)


@pytest.mark.parametrize("eig", [-2.0 - 1e-3, -3.1, -7.0])
@pytest.mark.parametrize("n", [6, 32])
def test_zsolve_periodic_vs_dense(n, eig):
    """Sherman-Morrison Thomas (src/triDiag.cpp:1961-2108) solves the cyclic system."""
    rng = np.random.default_rng(5)
    r = rng.standard_normal(n)
    A = np.diag(np.full(n, eig)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    A[0, -1] = A[-1, 0] = 1.0
    assert rel(oracle.zsolve_line(r, eig, "PP"), np.linalg.solve(A, r)) < 1e-11


def lap1d(n, h, lo, hi):
    """1D cell-centred second difference with ghost cells (P wrap, N mirror, D antisym)."""
    L = np.diag(np.full(n, -2.0)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    if lo == "P":
        L[0, -1] = L[-1, 0] = 1.0
    else:
        L[0, 0] += 1.0 if lo == "N" else -1.0
        L[-1, -1] += 1.0 if hi == "N" else -1.0
    return L / h ** 2


def dense_operator(shape, bc, box):
    nz, ny, nx = shape
    hx, hy, hz = (box[1] - box[0]) / nx, (box[3] - box[2]) / ny, (box[5] - box[4]) / nz
    Lx, Ly, Lz = lap1d(nx, hx, bc[0], bc[1]), lap1d(ny, hy, bc[2], bc[3]), lap1d(nz, hz, bc[4], bc[5])
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    # x fastest: index = i + nx*(j + ny*k)
    A = np.kron(Iz, np.kron(Iy, Lx)) + np.kron(Iz, np.kron(Ly, Ix)) + np.kron(Lz, np.kron(Iy, Ix))
    return A * hz ** 2


NONSING = ["DDDDDD", "DDNNDD", "NNDDNN", "DDDDNN", "PPPPDD", "DDPPDD", "NNPPDD", "DDNDDD",
           "NDDNDN", "DNNDND", "DDDDPP", "NDNDDD"]


@pytest.mark.parametrize("bc", NONSING)
def test_solve_vs_dense_operator(bc):
    """Non-singular combinations: the whole solve equals a dense solve of the 7-point
    operator whose right-hand side is rhs*dz^2 minus the Dirichlet boundary term."""
    shape = (5, 6, 8)
    box = [0.0, 1.0, 0.0, 1.5, -0.5, 0.5]
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, shape)
    A = dense_operator(shape, bc, box)
    fe = oracle.effective_rhs(f, bc, box)
    ref = np.linalg.solve(A, fe.ravel()).reshape(shape)
    for zero_mode in (0, 1):
        if zero_mode == 0 and bc[:4] in ("PPPP", "NNNN", "PPNN", "NNPP") and bc[4] != "P":
            continue  # reference rule changes the (0,0) mode when z has a Dirichlet face
        u = oracle.solve(f, bc, box, zero_mode=zero_mode)
        assert rel(u, ref) < 1e-11, (bc, zero_mode)
    # the oracle's own operator agrees with the dense one
    assert rel(oracle.apply_operator(ref, bc, box).ravel(), A @ ref.ravel()) < 1e-12


@pytest.mark.parametrize("bc", ["PPPPNN", "NNNNNN", "NNPPNN"])
def test_singular_mode_row0_rule(bc):
    """(0,0) column under x/y in {P,N} and z NN (src/triDiag.cpp:1753-1776,1629-1687):
    row 0 of that column becomes u0 = r0.  In physical terms: every non-(0,0) component and
    every plane z >= 1 of the xy-mean satisfy L u = f_eff exactly, and the xy-mean of u in
    plane 0 equals the xy-mean of f_eff there."""
    shape = (9, 8, 12)
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, shape)
    fe = oracle.effective_rhs(f, bc)
    u = oracle.solve(f, bc)
    r = fe - oracle.apply_operator(u, bc)
    scale = np.abs(fe).max() + 12 * np.abs(u).max()
    rm = r.mean(axis=(1, 2))
    assert np.abs(r - rm[:, None, None]).max() < 1e-13 * scale
    assert np.abs(rm[1:]).max() < 1e-13 * scale
    assert abs(u[0].mean() - fe[0].mean()) < 1e-13 * scale


def test_zero_mode_exact_with_dirichlet_z():
    """zero_mode=1 solves the true (non-singular) system when z has a Dirichlet face."""
    shape = (7, 8, 8)
    rng = np.random.default_rng(4)
    f = rng.uniform(-1, 1, shape)
    for bc in ("PPPPDD", "NNNNDN", "NNPPND"):
        ref = np.linalg.solve(dense_operator(shape, bc, [0, 1, 0, 1, 0, 1]),
                              oracle.effective_rhs(f, bc).ravel()).reshape(shape)
        assert rel(oracle.solve(f, bc, zero_mode=1), ref) < 1e-11
        assert rel(oracle.solve(f, bc, zero_mode=0), ref) > 1e-6  # reference rule differs


def grid(n):
    c = (np.arange(n) + 0.5) / n
    return np.meshgrid(c, c, c, indexing="ij")  # returns (z, y, x) order below


MMS = {
    # bc: (u(x,y,z), -lap u / u)
    "PPPPNN": (lambda x, y, z: np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.cos(np.pi * z), 9 * np.pi ** 2),
    "DDDDDD": (lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z), 3 * np.pi ** 2),
    "NNNNDD": (lambda x, y, z: np.cos(np.pi * x) * np.cos(2 * np.pi * y) * np.sin(np.pi * z), 6 * np.pi ** 2),
    "NDDDDD": (lambda x, y, z: np.cos(np.pi * x / 2) * np.sin(np.pi * y) * np.sin(np.pi * z), 2.25 * np.pi ** 2),
    "DDDNDD": (lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y / 2) * np.sin(np.pi * z), 2.25 * np.pi ** 2),
}


@pytest.mark.parametrize("bc", list(MMS))
def test_mms_second_order(bc):
    uex, lam = MMS[bc]
    errs = []
    for n in (16, 32, 64):
        z, y, x = grid(n)
        u = uex(x, y, z)
        f = -lam * u
        got = oracle.solve(f, bc, zero_mode=1)
        if bc == "PPPPNN":
            got = got - got.mean() + u.mean()  # null space of the all-Neumann/periodic column
        errs.append(np.sqrt(np.mean((got - u) ** 2)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4, (bc, errs)


def test_config1_manufactured_64():
    """Config 1 of the benchmark set: 64^3, PPPPNN, f = -(8+1) pi^2 cos2pix cos2piy cospiz."""
    uex, lam = MMS["PPPPNN"]
    z, y, x = grid(64)
    u = uex(x, y, z)
    got = oracle.solve(-lam * u, "PPPPNN")
    got = got - got.mean() + u.mean()
    err = np.sqrt(np.mean((got - u) ** 2))
    assert err < 2e-3


@pytest.mark.parametrize("n", [4, 5, 6, 16, 33, 64])
def test_singular_periodic_column_vs_dense(n):
    """Singular periodic-z column (eig = -2, shermanMorrisonThomasV1, src/triDiag.cpp:2112-2255):
    the answer is the cyclic system whose row 0 is replaced by
    u_0 + u_{N-1} = (r_0 + r_{N-1}) / 2, rows 1..N-1 unchanged (a dense LAPACK solve)."""
    rng = np.random.default_rng(n)
    r = rng.standard_normal(n)
    A = np.diag(np.full(n, -2.0)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    A[0, -1] = A[-1, 0] = 1.0
    A[0, :] = 0.0
    A[0, 0] = A[0, -1] = 1.0
    b = r.copy()
    b[0] = 0.5 * (r[0] + r[-1])
    assert rel(oracle.zsolve_line(r, -2.0, "PP"), np.linalg.solve(A, b)) < 1e-13


@pytest.mark.parametrize("bc", ["PPPPPP", "NNNNPP", "NNPPPP"])
def test_singular_periodic_column_physical(bc):
    """The same rule seen on the whole solve: every component except the xy-mean satisfies
    L u = f_eff exactly, the xy-mean satisfies its cyclic z equation on planes 1..N-1, and on
    plane 0 the xy-means obey u_0 + u_{N-1} = (f_0 + f_{N-1}) / 2."""
    shape = (10, 8, 12)
    rng = np.random.default_rng(3)
    f = rng.uniform(-1, 1, shape)
    fe = oracle.effective_rhs(f, bc)
    u = oracle.solve(f, bc)
    r = fe - oracle.apply_operator(u, bc)
    scale = np.abs(fe).max() + 12 * np.abs(u).max()
    rm = r.mean(axis=(1, 2))
    assert np.abs(r - rm[:, None, None]).max() < 1e-13 * scale
    assert np.abs(rm[1:]).max() < 1e-13 * scale
    um, fm = u.mean(axis=(1, 2)), fe.mean(axis=(1, 2))
    assert abs(um[0] + um[-1] - 0.5 * (fm[0] + fm[-1])) < 1e-13 * scale


def reference_mms(bc, box, n):
    """The reference's built-in manufactured solution (initializeTrigonometric / exactValue,
    src/pencilDcmp.cpp:1837-1866,1871-2035): u = prod_d g_d(omega_d x_d) with
    omega = (1, 2, 3) pi and g = sin on Dirichlet axes, cos on Neumann axes (z: sin), and
    f = -|omega|^2 u.  Its Dirichlet face values feed modifyRhsDirichlet."""
    om = np.pi * np.array([1.0, 2.0, 3.0])
    c = [box[2 * a] + (np.arange(n) + 0.5) * (box[2 * a + 1] - box[2 * a]) / n for a in range(3)]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    gx = np.sin if bc[0] == "D" else np.cos
    gy = np.sin if bc[2] == "D" else np.cos
    u = gx(om[0] * x) * gy(om[1] * y) * np.sin(om[2] * z)
    return u, -(om ** 2).sum() * u


@pytest.mark.parametrize("bc,box", [
    ("DDDDDD", [-0.3, 0.7, 0.2, 1.2, -0.3, 0.7]),
    ("NNNNDD", [0.0, 1.0, 0.0, 1.0, -0.3, 0.7]),
    ("DDNNDD", [-0.3, 0.7, 0.0, 1.0, -0.3, 0.7]),
])
def test_dirichlet_term_second_order(bc, box):
    """Dirichlet boundary term (modifyRhsDirichlet, src/pencilDcmp.cpp:4623-4835): on boxes
    shifted so the face values of the manufactured solution are O(1) (dx = dy = dz, as the
    reference's term assumes), the solve converges to the exact solution at second order.
    A wrong sign or factor on the term leaves an O(1) error that does not shrink."""
    errs = []
    for n in (16, 32, 64):
        u, f = reference_mms(bc, box, n)
        faces = [u[:, :, 0], u[:, :, -1], u[:, 0, :], u[:, -1, :], u[0], u[-1]]
        assert max(np.abs(g).max() for g, c in zip(faces, bc) if c == "D") > 0.3
        got = oracle.solve(f, bc, box, zero_mode=1)
        errs.append(np.sqrt(np.mean((got - u) ** 2)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.8 < r1 < 4.2 and 3.8 < r2 < 4.2, (bc, errs)
    assert errs[-1] < 1e-3
