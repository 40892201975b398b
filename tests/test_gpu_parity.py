"""GPU parity of the FP64 solve path against the CPU oracle (through the C ABI).

Bar: max-norm difference <= 1e-12 relative to max|u_oracle| (the FP64 parity bar), on
seeded U[-1,1] right-hand sides, for every boundary combination the oracle restates.
Combinations the reference's algorithm cannot represent (periodic x with non-periodic y)
are checked through the discrete residual of the 7-point operator instead.
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

TOL = 1e-12

SHAPES = [(10, 16, 32), (20, 24, 48), (33, 64, 16), (7, 16, 96)]
# (x, y, z) faces; every combination below is solvable by the reference's algorithm
# except the quarter-wave ND/DN axes (extension), which the oracle restates from definition.
COMBOS = ["PPPPNN", "NNNNDD", "DDDDDD", "DDNNDD", "NNDDNN", "DDDDNN", "NNNNNN",
          "PPPPDD", "DDPPDD", "NNPPNN", "DDNDDD", "NDDNDN", "DNNNDD", "NDNDNN"]


def rel_maxdiff(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_selftest_lane_rotation():
    print("selftest rot16", flush=True)
    assert pa.selftest() == 0


@pytest.mark.parametrize("zsolver", [0, 1, 2, 3, 4], ids=["pcr", "scan", "mg", "pcrfull", "cr"])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", COMBOS)
def test_parity_fp64(bc, shape, zsolver):
    print(f"parity {bc} {shape} zsolver={zsolver}", flush=True)
    rng = np.random.default_rng(sum(map(ord, bc)) * 1000 + shape[0] * 37 + shape[2])
    f = rng.uniform(-1, 1, shape)
    box = [0, 1, 0, 1, 0, 1]
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
        got = pa.solve(f, bc, box, zero_mode=zero_mode, zsolver=zsolver)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} {shape} zero_mode={zero_mode}: rel max diff {err:.3e}"


# (zsolver, zlayout): PCR -> Thomas from 16-column z tiles (half a tile per 8-column block)
# and in place; the alternates in place
LONG_CASES = [(0, 1), (0, 0), (2, 1), (3, 1), (4, 1)]


@pytest.mark.parametrize("case", LONG_CASES, ids=["pcr-tiles", "pcr-inplace", "mg", "pcrfull", "cr"])
@pytest.mark.parametrize("nz", [513, 530, 640, 777, 1000, 1024])
def test_parity_long_columns(nz, case):
    """FP64 columns of 513..1024 rows (33..64 rows per lane for PCR -> Thomas, 8-column blocks
    with Thomas factors in LDS); 20 columns leave the last block half empty"""
    zsolver, zlayout = case
    print(f"long column {nz} zsolver={zsolver} zlayout={zlayout}", flush=True)
    rng = np.random.default_rng(nz + 5)
    shape = (nz, 12, 20)
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    f = rng.uniform(-1, 1, shape)
    for bc in ("NNNNDD", "PPPPNN", "DDNDDD", "DDDDDN"):
        for zero_mode in (0, 1):
            ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
            got = pa.solve(f, bc, box, zero_mode=zero_mode, zsolver=zsolver, zlayout=zlayout)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"{bc} nz={nz} zero_mode={zero_mode}: rel max diff {err:.3e}"


@pytest.mark.parametrize("shape", [(530, 640, 20), (1024, 768, 8), (500, 768, 20), (512, 1024, 16)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zlayout", [1, 0], ids=["tiles", "inplace"])
def test_parity_ztiles_from_8col_y(shape, zlayout):
    """y lines of 640 to 1024 points run 8 columns per block, so two y blocks fill each
    16-column z tile; z blocks take whole tiles (up to 512 rows) or half (longer columns)"""
    print(f"z tiles from 8-column y blocks {shape} zlayout={zlayout}", flush=True)
    rng = np.random.default_rng(shape[0] + shape[1])
    f = rng.uniform(-1, 1, shape)
    box = [0.0, shape[2] / 64.0, 0.0, shape[1] / 64.0, 0.0, shape[0] / 64.0]
    for bc in ("DDNDDD", "PPPPNN"):
        ref = oracle.solve(f, bc, box, zero_mode=1)
        got = pa.solve(f, bc, box, zero_mode=1, zlayout=zlayout)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} {shape}: rel max diff {err:.3e}"


def test_long_columns_limits_and_autotune():
    """partitioned Thomas stops at 512 rows and the tuner skips it beyond; a wide 1024-row
    solve through the tuned handle"""
    print("long column limits", flush=True)
    rng = np.random.default_rng(3)
    f = rng.uniform(-1, 1, (600, 16, 16))
    with pytest.raises(pa.PittPackError) as e:
        pa.solve(f, "NNNNDD", zsolver=1)
    assert e.value.code == pa.UNSUPPORTED_SIZE
    # dx = dy = dz as in the benchmark grids (a unit box would make dz 10x finer than dx and
    # the low modes' columns ill-conditioned enough to amplify rounding past the bar)
    box = [0, 96 / 1024, 0, 64 / 1024, 0, 1]
    f = rng.uniform(-1, 1, (1024, 64, 96))
    ref = oracle.solve(f, "DDNDDD", box)
    with pa.PoissonHIP(96, 64, 1024) as M:
        M.set_option(pa.OPT_AUTOTUNE, 1)
        M.assignBoundary("DDNDDD")
        M.setBox(box)
        M.setCoords(2)
        for _ in range(2):
            M.assignRhs(f)
            M.pittPack()
            err = rel_maxdiff(M.getSolution(), ref)
            assert err <= TOL, f"rel max diff {err:.3e}"
        tuned = M.tuning()
    assert tuned["tuned"] and tuned["zsolver"] == 0


# line lengths beyond the power-of-two/radix-3 set: the reference's small fixture sizes
# 16 x 12 x 10 and 24 x 20 x 18 (N = 12, 20), the shortest lines (N = 8, 12), radix-5
# lengths up to N = 640 and radix-7 lengths up to N = 896 (shape (nz, ny, nx))
LENGTH_SHAPES = [(10, 12, 16), (18, 20, 24), (12, 8, 40), (9, 80, 20), (5, 160, 12), (4, 40, 640), (4, 320, 8),
                 (7, 28, 56), (5, 448, 112), (4, 56, 896), (6, 224, 28)]


@pytest.mark.parametrize("shape", LENGTH_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", COMBOS)
def test_parity_line_lengths(bc, shape):
    print(f"line lengths {bc} {shape}", flush=True)
    rng = np.random.default_rng(sum(map(ord, bc)) * 7 + shape[1] * 13 + shape[2])
    f = rng.uniform(-1, 1, shape)
    box = [0, 1.5, 0, 1, 0, 2]
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
        got = pa.solve(f, bc, box, zero_mode=zero_mode)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} {shape} zero_mode={zero_mode}: rel max diff {err:.3e}"


@pytest.mark.parametrize("zsolver", [0, 1, 2, 3, 4], ids=["pcr", "scan", "mg", "pcrfull", "cr"])
@pytest.mark.parametrize("nz", [4, 5, 15, 16, 17, 31, 33, 48, 49, 64, 100, 128, 255, 256, 500, 511, 512])
def test_parity_z_lengths(nz, zsolver):
    """every z length class: J = 1..64 rows per lane, padded and full columns; uniform
    spacing dx = dy = dz as in the benchmark grids"""
    print(f"z length {nz} zsolver={zsolver}", flush=True)
    rng = np.random.default_rng(nz)
    shape = (nz, 16, 16)
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    f = rng.uniform(-1, 1, shape)
    for bc in ("NNNNDD", "PPPPNN", "DDDDDN"):
        ref = oracle.solve(f, bc, box, zero_mode=1)
        got = pa.solve(f, bc, box, zero_mode=1, zsolver=zsolver)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} nz={nz}: rel max diff {err:.3e}"


@pytest.mark.parametrize("bc", ["PPDDNN", "PPNNDD", "PPDNDD"])
def test_periodic_x_mixed_y_residual(bc):
    """x periodic with y non-periodic: not representable in the reference's algorithm
    (its DCT/DST steps drop the imaginary part); checked by the discrete residual."""
    print(f"residual {bc}", flush=True)
    rng = np.random.default_rng(7)
    shape = (12, 16, 32)
    f = rng.uniform(-1, 1, shape)
    u = pa.solve(f, bc)
    fe = oracle.effective_rhs(f, bc)
    r = fe - oracle.apply_operator(u, bc)
    scale = max(np.abs(fe).max(), 12.0 * np.abs(u).max())
    rmax = float(np.abs(r).max())
    assert rmax <= 1e-12 * scale, f"{bc}: residual {rmax:.3e} vs scale {scale:.3e}"


def test_nonunit_box_and_aspect():
    print("box", flush=True)
    rng = np.random.default_rng(3)
    f = rng.uniform(-1, 1, (24, 32, 64))
    box = [-1.0, 2.0, 0.5, 1.25, -0.3, 0.7]
    for bc in ("DDNNDD", "PPPPNN", "NDDDND"):
        ref = oracle.solve(f, bc, box)
        got = pa.solve(f, bc, box)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc}: rel max diff {err:.3e}"


def test_repeated_solves_reuse_tables():
    """flow-solver pattern: one handle, several right-hand sides"""
    print("repeat", flush=True)
    rng = np.random.default_rng(11)
    shape = (16, 32, 32)
    with pa.PoissonHIP(32, 32, 16) as M:
        M.assignBoundary("NNNNDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        for _ in range(3):
            f = rng.uniform(-1, 1, shape)
            M.assignRhs(f)
            M.pittPack()
            err = rel_maxdiff(M.getSolution(), oracle.solve(f, "NNNNDD", zero_mode=1))
            assert err <= TOL, f"repeat solve rel max diff {err:.3e}"


def test_call_order_and_errors():
    print("errors", flush=True)
    with pa.PoissonHIP(32, 32, 16) as M:
        with pytest.raises(pa.PittPackError) as e:
            M.pittPack()
        assert e.value.code == pa.CALL_ORDER
        with pytest.raises(pa.PittPackError) as e:
            M.assignBoundary("PNPPDD")
        assert e.value.code == pa.UNSUPPORTED_BC
    with pytest.raises(pa.PittPackError) as e:
        pa.PoissonHIP(1030, 32, 16)
    assert e.value.code == pa.UNSUPPORTED_SIZE
    with pytest.raises(pa.PittPackError) as e:
        pa.PoissonHIP(32, 32, 1025)
    assert e.value.code == pa.UNSUPPORTED_SIZE
    # FP64 columns: up to 1024 rows with the default z solver, 512 with partitioned Thomas
    with pa.PoissonHIP(32, 32, 513) as M:
        M.assignBoundary("NNNNDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        f = np.random.default_rng(2).uniform(-1, 1, (513, 32, 32))
        M.assignRhs(f)
        M.pittPack()
        assert rel_maxdiff(M.getSolution(), oracle.solve(f, "NNNNDD", zero_mode=1)) <= TOL
        M.set_option(pa.OPT_ZSOLVER, 1)
        M.assignRhs(np.zeros((513, 32, 32)))
        with pytest.raises(pa.PittPackError) as e:
            M.pittPack()
        assert e.value.code == pa.UNSUPPORTED_SIZE


def _solve_opts(f, bc, box, zthomas=None, zsolver=None, zlayout=None):
    nz, ny, nx = f.shape
    with pa.PoissonHIP(nx, ny, nz) as M:
        M.assignBoundary(bc)
        M.setBox(box)
        M.setCoords(2)
        if zsolver is not None:
            M.set_option(pa.OPT_ZSOLVER, zsolver)
        if zlayout is not None:
            M.set_option(pa.OPT_ZLAYOUT, zlayout)
        if zthomas is not None:
            M.set_option_double(pa.OPT_ZTHOMAS_TAU, zthomas)
        M.assignRhs(f)
        M.pittPack()
        return M.getSolution()


def test_zthomas_thin_cells():
    """cells six times thinner in z than in x and y with Neumann z faces: the low (kx, ky)
    columns have |eig + 2| ~ 1e-5, where the PCR steps lose about ten times more digits than
    the reference's Thomas (2e-12 against the oracle with PP_OPT_ZTHOMAS_TAU = 0); the default
    threshold hands those columns to the reference's Thomas sweeps and restores the FP64 bar"""
    print("zthomas thin cells NNNNNN (96, 64, 128)", flush=True)
    f = prng.uniform_pm1((96, 64, 128), seed=41)
    box = [0.0, 2.0, 0.0, 1.0, 0.0, 0.25]
    ref = oracle.solve(f, "NNNNNN", box, zero_mode=1)
    e_pcr = rel_maxdiff(_solve_opts(f, "NNNNNN", box, zthomas=0.0), ref)
    e_thomas = rel_maxdiff(_solve_opts(f, "NNNNNN", box), ref)
    print(f"  PCR {e_pcr:.2e}  with Thomas columns {e_thomas:.2e}", flush=True)
    assert e_thomas <= TOL and e_thomas <= e_pcr, (e_pcr, e_thomas)


@pytest.mark.parametrize("zsolver,zlayout", [(0, 1), (0, 0), (1, 0)])
@pytest.mark.parametrize("bc", ["DDNDDD", "NNNNDD", "PPPPNN", "NNNNNN"])
def test_zthomas_paths(bc, zsolver, zlayout):
    """the Thomas columns on z tiles and in place, after either direct z solver, both zero-mode
    rules; a threshold of 0.5 sends many columns there"""
    print(f"zthomas {bc} zsolver={zsolver} zlayout={zlayout}", flush=True)
    f = prng.uniform_pm1((40, 48, 64), seed=sum(map(ord, bc)) + zsolver)
    box = [0, 1, 0, 1, 0, 1]
    ref = oracle.solve(f, bc, box, zero_mode=1)
    got = _solve_opts(f, bc, box, zthomas=0.5, zsolver=zsolver, zlayout=zlayout)
    assert rel_maxdiff(got, ref) <= TOL


# periodic z: circulant columns by z FFT on the GPU, the singular column by the reference's
# own branch; the oracle follows shermanMorrisonThomas[V1]
ZPER_COMBOS = ["PPPPPP", "NNNNPP", "DDDDPP", "DDNNPP", "NDDNPP", "NNPPPP"]


@pytest.mark.parametrize("nz", [8, 10, 14, 16, 20, 24, 40, 48, 56, 64, 96, 128, 320, 448, 512])
@pytest.mark.parametrize("bc", ZPER_COMBOS)
def test_parity_periodic_z(bc, nz):
    print(f"periodic z {bc} nz={nz}", flush=True)
    rng = np.random.default_rng(nz * 7 + len(bc))
    shape = (nz, 16, 32)
    f = rng.uniform(-1, 1, shape)
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    ref = oracle.solve(f, bc, box)
    got = pa.solve(f, bc, box)
    err = rel_maxdiff(got, ref)
    assert err <= TOL, f"{bc} nz={nz}: rel max diff {err:.3e}"


def test_periodic_z_residual():
    """periodic x with non-periodic y and periodic z: the discrete residual, FP64 and mixed"""
    print("periodic z residual", flush=True)
    rng = np.random.default_rng(17)
    f = rng.uniform(-1, 1, (32, 16, 32))
    u = pa.solve(f, "PPDDPP")
    fe = oracle.effective_rhs(f, "PPDDPP")
    r = fe - oracle.apply_operator(u, "PPDDPP")
    assert float(np.abs(r).max()) <= 1e-12 * max(np.abs(fe).max(), 12.0 * np.abs(u).max())
    um = pa.solve(f, "PPDDPP", precision=1)
    assert rel_maxdiff(um, u) <= TOL


# line lengths with an alternative launch shape (radix plan / block) from 16 to 512
VARIANT_SHAPES = [(10, 16, 32), (20, 24, 48), (6, 128, 256), (4, 64, 512), (5, 256, 128), (7, 80, 80)]


@pytest.mark.parametrize("shape", VARIANT_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", ["PPPPNN", "NNNNDD", "DDNDDD", "DDDDNN", "NDDNDN"])
def test_parity_line_variant(bc, shape):
    """every FP64 line stage in its alternative launch shape (the autotuner's other choice)"""
    print(f"line variant {bc} {shape}", flush=True)
    rng = np.random.default_rng(sum(map(ord, bc)) + shape[2])
    f = rng.uniform(-1, 1, shape)
    box = [0, 1, 0, 1, 0, 1]
    ref = oracle.solve(f, bc, box, zero_mode=1)
    got = pa.solve(f, bc, box, zero_mode=1, line_variant=1)
    err = rel_maxdiff(got, ref)
    assert err <= TOL, f"{bc} {shape}: rel max diff {err:.3e}"


def test_autotune_choice_and_parity():
    """the autotuner times the candidates, keeps a configuration, and the solve still meets
    the FP64 bar; a second handle of the same shape reuses the cached choice"""
    print("autotune 64x96x128", flush=True)
    rng = np.random.default_rng(7)
    shape = (64, 96, 128)
    f = rng.uniform(-1, 1, shape)
    box = [0, 1, 0, 1, 0, 1]
    ref = oracle.solve(f, "DDNDDD", box, zero_mode=1)
    for _ in range(2):
        with pa.PoissonHIP(shape[2], shape[1], shape[0]) as M:
            M.set_option(pa.OPT_AUTOTUNE, 1)
            M.assignBoundary("DDNDDD")
            M.setBox(box)
            M.setCoords(2)
            M.assignRhs(f)
            M.pittPack()
            got = M.getSolution()
            t = M.tuning()
        assert t["tuned"], t
        assert t["zsolver"] in (0, 1)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"autotuned {t}: rel max diff {err:.3e}"
