"""Second-order convergence of the HIP path on manufactured solutions.

For u(x, y, z) a product of eigenfunctions that satisfies the faces' conditions exactly, the
7-point discretisation's error against u falls by 4 per halving of the spacing.  The FP64
form and the mixed form (FP32 work + two FP64-residual refinement steps) must both show it
(RMS error ratio in (3.6, 4.4) over 32^3 -> 64^3 -> 128^3), with the error of each grid
itself within 1e-12 relative of the oracle's.  Config 1 of the benchmark set (64^3, PPPPNN,
f = -(8+1) pi^2 cos 2 pi x cos 2 pi y cos pi z) is checked as specified.
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa

pytestmark = pytest.mark.gpu


def grid(n):
    c = (np.arange(n) + 0.5) / n
    return np.meshgrid(c, c, c, indexing="ij")  # (z, y, x)


MMS = {
    # bc: (u(x, y, z), -lap u / u)
    "PPPPNN": (lambda x, y, z: np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y) * np.cos(np.pi * z), 9 * np.pi ** 2),
    "DDDDDD": (lambda x, y, z: np.sin(np.pi * x) * np.sin(np.pi * y) * np.sin(np.pi * z), 3 * np.pi ** 2),
    "NNNNDD": (lambda x, y, z: np.cos(np.pi * x) * np.cos(2 * np.pi * y) * np.sin(np.pi * z), 6 * np.pi ** 2),
    # the benchmark's boundary set: x D-D, y N-D (cos(pi y / 2): zero slope at 0, zero at 1), z D-D
    "DDNDDD": (lambda x, y, z: np.sin(np.pi * x) * np.cos(np.pi * y / 2) * np.sin(np.pi * z), 2.25 * np.pi ** 2),
    "NDDDNN": (lambda x, y, z: np.cos(np.pi * x / 2) * np.sin(np.pi * y) * np.cos(np.pi * z), 2.25 * np.pi ** 2),
}


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("bc", list(MMS))
def test_mms_second_order_gpu(bc, precision):
    print(f"mms {bc} precision={precision}", flush=True)
    uex, lam = MMS[bc]
    errs = []
    for n in (32, 64, 128):
        z, y, x = grid(n)
        u = uex(x, y, z)
        f = -lam * u
        got = pa.solve(f, bc, precision=precision)
        if n <= 64:
            ref = oracle.solve(f, bc, zero_mode=1)
            d = float(np.abs(got - ref).max() / np.abs(ref).max())
            assert d <= 1e-12, f"{bc} n={n}: rel max diff vs oracle {d:.2e}"
        if bc == "PPPPNN":
            got = got - got.mean() + u.mean()  # null space of the periodic / Neumann column
        errs.append(float(np.sqrt(np.mean((got - u) ** 2))))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4, (bc, precision, errs)


def test_config1_manufactured_64_gpu():
    """64^3 PPPPNN with the manufactured right-hand side of benchmark config 1"""
    print("config 1 manufactured", flush=True)
    uex, lam = MMS["PPPPNN"]
    z, y, x = grid(64)
    u = uex(x, y, z)
    f = -lam * u
    got = pa.solve(f, "PPPPNN")
    ref = oracle.solve(f, "PPPPNN")
    assert float(np.abs(got - ref).max() / np.abs(ref).max()) <= 1e-12
    got = got - got.mean() + u.mean()
    assert float(np.sqrt(np.mean((got - u) ** 2))) < 2e-3
