"""The reference's built-in verification case on the GPU: initializeTrigonometric, getError and
fillTrigonometric (src/pencilDcmp.cpp:1871-2208,5685-5765) through the C ABI.

* the right-hand side pp_init_trigonometric writes equals the reference's formulas (restated
  here in numpy) at rounding level, for every case the reference distinguishes;
* the solve of that right-hand side meets the 1e-12 bar against the oracle;
* pp_get_error equals the RMS difference to the exact solution computed in numpy;
* on boundary conditions where the manufactured solution satisfies the boundary conditions
  (including main.cpp's NNNNDD), pp_get_error falls at second order (ratio ~4 per halving)."""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa

pytestmark = pytest.mark.gpu

PI = np.pi


def exact_value(omega, x, tag):
    """exactValue (src/pencilDcmp.cpp:1837-1866)"""
    if abs(omega) < 1e-6:
        return np.ones_like(x)
    return np.cos(omega * x) if tag == 1 else np.sin(omega * x)


def tags(bc):
    """extractTag (src/pencilDcmp.cpp:3636-3668) for x, y (z: -1 D, -2 N)"""
    t = []
    for a in range(2):
        lo, hi = bc[2 * a], bc[2 * a + 1]
        t.append(0 if lo == hi == "D" else 1 if lo == hi == "N" else 2 if lo == hi == "P" else -1)
    t.append(-1 if bc[4] == "D" else -2 if bc[4] == "N" else 2)
    return t


def manufactured(bc, box, shape, part):
    nz, ny, nx = shape
    c = [box[2 * a] + (np.arange(n) + 0.5) * (box[2 * a + 1] - box[2 * a]) / n for a, n in enumerate((nx, ny, nz))]
    z, y, x = np.meshgrid(c[2], c[1], c[0], indexing="ij")
    w = PI * np.array([1.0, 2.0, 3.0])
    if bc[0] == "P" and bc[2] == "P" and bc[4] == "P":
        u = (np.sin if part else np.cos)(w[1] * (x + y + z))
        return u, -3 * w[1] ** 2 * u
    if bc[0] == "P" or bc[2] == "P":
        u = np.sin(w[1] * z) * (np.sin if part else np.cos)(w[1] * (x + y))
        return u, -3 * w[1] ** 2 * u
    t = tags(bc)
    u = np.zeros(shape) if part else exact_value(w[0], x, t[0]) * exact_value(w[1], y, t[1]) * exact_value(w[2], z, t[2])
    return u, -(w ** 2).sum() * u


def run(bc, box, shape, part):
    nz, ny, nx = shape
    M = pa.PoissonHIP(nx, ny, nz)
    M.assignBoundary(bc)
    M.setBox(box)
    M.setCoords(2)
    M.initializeTrigonometric(part)
    f = M.getSolution()  # the field holds the right-hand side until the solve
    M.pittPack()
    return M, f, M.getSolution(), M.getError()


CASES = [("NNNNDD", [0, 1, 0, 1, 0, 1]), ("DDDDDD", [-0.3, 0.7, 0.2, 1.2, -0.3, 0.7]),
         ("PPPPDD", [0, 1, 0, 1, 0, 1]), ("PPPPPP", [0, 1, 0, 1, 0, 1]), ("DDNNDD", [0, 1, 0, 1, 0, 1]),
         ("NNPPNN", [0, 1, 0, 1, 0, 1]), ("DDNDDD", [0, 1, 0, 1, 0, 1])]


@pytest.mark.parametrize("part", [0, 1])
@pytest.mark.parametrize("bc,box", CASES)
def test_trigonometric_rhs_solution_and_error(bc, box, part):
    shape = (10, 12, 16)
    print(f"trigonometric {bc} part {part}", flush=True)
    M, f, u, err = run(bc, box, shape, part)
    ue, fe = manufactured(bc, box, shape, part)
    assert np.abs(f - fe).max() <= 1e-13 * max(np.abs(fe).max(), 1.0)
    ref = oracle.solve(f, bc, box, zero_mode=1)
    assert np.abs(u - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300)
    want = np.sqrt(np.mean((u - ue) ** 2))
    assert abs(err - want) <= 1e-12 * max(want, 1e-300), (err, want)


@pytest.mark.parametrize("bc,box,part", [("NNNNDD", [0, 1, 0, 1, 0, 1], 0), ("DDDDDD", [-0.3, 0.7, 0.2, 1.2, -0.3, 0.7], 0),
                                         ("PPPPDD", [0, 1, 0, 1, 0, 1], 0), ("PPPPDD", [0, 1, 0, 1, 0, 1], 1),
                                         ("PPPPPP", [0, 1, 0, 1, 0, 1], 0), ("PPPPPP", [0, 1, 0, 1, 0, 1], 1)])
def test_get_error_second_order(bc, box, part):
    print(f"convergence {bc} part {part}", flush=True)
    errs = []
    for n in (16, 32, 64):
        M, _, _, err = run(bc, box, (n, n, n), part)
        M.close()
        errs.append(err)
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.7 < r1 < 4.3 and 3.7 < r2 < 4.3, (bc, errs)
    assert errs[-1] < 2e-3


def test_fill_trigonometric_and_error_of_user_rhs():
    """INITANALYTIC = 0 path: main.cpp fills -4 pi^2 cos(2 pi z) and getError compares with
    cos(2 pi z); here on PPPPDD shifted so the z faces are where cos(2 pi z) = 0."""
    nx, ny, nz = 8, 8, 32
    box = [0, 1, 0, 1, 0.25, 0.75]
    M = pa.PoissonHIP(nx, ny, nz)
    M.assignBoundary("PPPPDD")
    M.setBox(box)
    M.setCoords(2)
    f = M.fillTrigonometric()
    z = 0.25 + (np.arange(nz) + 0.5) * 0.5 / nz
    assert np.abs(f - (-4 * PI ** 2 * np.cos(2 * PI * z))[:, None, None]).max() < 1e-12
    M.assignRhs(f)
    M.pittPack()
    u = M.getSolution()
    want = np.sqrt(np.mean((u - np.cos(2 * PI * z)[:, None, None]) ** 2))
    err = M.getError()
    assert abs(err - want) <= 1e-12 * want
    assert err < 2e-3
    # an initialised case stays selected until a new right-hand side is assigned
    M.initializeTrigonometric(0)
    M.pittPack()
    e0 = M.getError()
    M.assignRhs(f)
    M.pittPack()
    assert M.getError() == err and e0 != err


def test_trigonometric_call_order():
    M = pa.PoissonHIP(8, 8, 8)
    with pytest.raises(pa.PittPackError) as e:
        M.initializeTrigonometric(0)
    assert e.value.code == pa.CALL_ORDER
    M.assignBoundary("NNNNDD")
    M.setBox([0, 1, 0, 1, 0, 1])
    M.setCoords(2)
    with pytest.raises(pa.PittPackError) as e:
        M.initializeTrigonometric(2)
    assert e.value.code == pa.INPUT_ARGS
