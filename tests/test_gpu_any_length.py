"""Line lengths without a compiled Stockham plan, on the GPU against the oracle.

Odd lengths and lengths whose half has a prime factor above 7 (N = 9, 33, 37, 97, 101, 211,
257, 509, ...) run through the arbitrary-length transforms (Bluestein convolution through
the Stockham engine); the reference takes any length from FFTW / cuFFT and SignalProc
handles odd sizes (src/signalProc.cpp:34-44,161-195).  Every transform kind appears in x and
in y, in the FP64, mixed and batched forms, and periodic z of any length.

Bar: max-norm difference <= 1e-12 relative to max|u_oracle| (FP64 and, after its two
refinement steps, the mixed form); batched members bitwise equal to the single-solve handle.
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

TOL = 1e-12
LENGTHS = [9, 33, 37, 97, 101, 211, 257, 509]
# every transform kind in x (PP, NN, DD, ND, DN) and in y
COMBOS = ["PPPPNN", "NNNNDD", "DDDDDD", "DDNDDD", "NDDNDN", "DNNNDD", "NNDDNN", "DDPPDD", "NNPPDD"]


def rel_maxdiff(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_every_length_supported():
    for n in range(2, 1025):
        assert pa.size_supported(n), n
    assert not pa.size_supported(1) and not pa.size_supported(1025)


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("bc", COMBOS)
def test_parity_any_length_fp64(bc, n, axis):
    shape = (6, 12, n) if axis == "x" else (5, n, 20)
    print(f"any length fp64 {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=n * 11 + len(axis) + sum(map(ord, bc)))
    box = [0, 1, 0, 1.5, 0, 0.5]
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
        got = pa.solve(f, bc, box, zero_mode=zero_mode)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} {shape} zero_mode={zero_mode}: rel max diff {err:.3e}"


# both axes arbitrary, lengths over 512 (convolution length 2048, four waves per line) and
# the tiny ones (N = 2 .. 7)
BOTH_SHAPES = [(7, 37, 33), (4, 101, 211), (4, 600, 24), (4, 16, 1000), (4, 1023, 8), (5, 3, 5), (4, 7, 2),
               (6, 30, 22), (9, 44, 100)]


@pytest.mark.parametrize("shape", BOTH_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", COMBOS)
def test_parity_any_length_both_axes(bc, shape):
    print(f"any length both axes {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=shape[1] * 3 + shape[2])
    box = [0, 1, 0, 1, 0, 1]
    ref = oracle.solve(f, bc, box, zero_mode=1)
    got = pa.solve(f, bc, box, zero_mode=1)
    err = rel_maxdiff(got, ref)
    assert err <= TOL, f"{bc} {shape}: rel max diff {err:.3e}"


@pytest.mark.parametrize("n", LENGTHS + [600, 1000])
@pytest.mark.parametrize("bc", ["PPPPNN", "NNNNDD", "DDNDDD", "NDDNDN", "DNNNDD", "NNNNNN"])
def test_parity_any_length_mixed(bc, n):
    """FP32-stored arbitrary-length transforms with FP64 refinement: the FP64 bar after two
    steps.  The box makes dz equal to the long axis' spacing: the refinement converges to the
    exact 7-point solution, which on cells much thinner than dz differs from the reference's
    spectral solution by its eigenvalue rounding (test_gpu_mixed.py,
    test_mixed_converges_to_stencil_solution)."""
    shape = (5, 24, n) if n % 2 else (4, n, 24)
    nz = shape[0]
    box = [0, 1, 0, 1, 0, nz / n]
    # With Neumann z faces the columns of the lowest (kx, ky) modes have eig + 2 ~ (pi / n)^2,
    # and the reference's eigenvalue arithmetic (-2 + 2 cos theta, then -2 + ... sums), which
    # the oracle and the FP64 form share, carries a relative error of about 4 eps / (pi / n)^2
    # there; the refinement converges to the exact 7-point solution instead.  The bar is the
    # larger of 1e-12 and that bound (Dirichlet z lifts every column: 1e-12).
    tol = TOL if "D" in bc[4:] else max(TOL, 4 * np.finfo(float).eps * (n / np.pi) ** 2)
    print(f"any length mixed {bc} {shape} tol {tol:.1e}", flush=True)
    f = prng.uniform_pm1(shape, seed=n + 3)
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
        got = pa.solve(f, bc, box, zero_mode=zero_mode, precision=1)
        err = rel_maxdiff(got, ref)
        assert err <= tol, f"{bc} {shape} zero_mode={zero_mode}: rel max diff {err:.3e}"


ZPER_COMBOS = ["PPPPPP", "NNNNPP", "DDDDPP", "NDDNPP", "NNPPPP"]


@pytest.mark.parametrize("nz", [5, 9, 22, 33, 37, 101, 257, 600])
@pytest.mark.parametrize("bc", ZPER_COMBOS)
def test_parity_periodic_z_any_length(bc, nz):
    """periodic z lengths without a compiled complex plan: real z-line transforms around a
    division by the circulant eigenvalues; the singular column by the reference's branch"""
    shape = (nz, 12, 20)
    print(f"periodic z any length {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=nz * 5 + len(bc))
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    ref = oracle.solve(f, bc, box)
    got = pa.solve(f, bc, box)
    err = rel_maxdiff(got, ref)
    assert err <= TOL, f"{bc} nz={nz}: rel max diff {err:.3e}"


ANY_MEMBERS = [
    (33, 37, 17, "PPPPNN", [0, 1, 0, 1, 0, 1]),
    (97, 48, 40, "NNNNDD", [0, 2, 0, 1, 0, 1]),
    (48, 101, 33, "DDNDDD", [0, 1, 0, 1, 0, 1]),
    (211, 9, 16, "NDDNDN", [0, 1, 0, 1, 0, 1]),
    (20, 257, 9, "DNNNDD", [0, 1, 0, 1, 0, 1]),
    (509, 24, 5, "NNNNNN", [0, 1, 0, 1, 0, 1]),
    (32, 48, 17, "DDDDDD", [0, 1, 0, 1, 0, 1]),
    (30, 300, 6, "NNPPDD", [0, 1, 0, 1, 0, 1]),
]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_any_length(precision):
    """arbitrary lengths beside compiled ones in one batch: each member equals the
    single-solve handle bitwise and meets the bar against the oracle"""
    print(f"batch any length precision={precision}", flush=True)
    with pa.Batch(precision=precision) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in ANY_MEMBERS]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, ANY_MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=300 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, ANY_MEMBERS):
            got = B.get_solution(i)
            single = pa.solve(rhs[i], bc, box, precision=precision)
            assert np.array_equal(got, single), f"member {i} {bc}: batched != single ({rel_maxdiff(got, single):.2e})"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"member {i} {bc} {(nx, ny, nz)}: {err:.3e}"
