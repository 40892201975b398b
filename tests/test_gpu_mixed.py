"""GPU parity of the mixed-precision form against the CPU oracle (through the C ABI).

The mixed form runs the transforms and z solves in FP32 (FP64 twiddles and eigenvalues),
solves the near-singular columns (|eig+2| < tau) and the pinned (0,0) column in FP64, then
refines with the FP64 residual of the 7-point operator.  After the default two refinement
steps the result must meet the FP64 bar: max-norm difference <= 1e-12 relative to
max|u_oracle|.  Without refinement it is an FP32 solve (checked at 1e-5).
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

TOL = 1e-12

SHAPES = [(10, 16, 32), (33, 64, 16), (7, 16, 96), (18, 20, 24), (6, 160, 40), (5, 12, 320), (9, 56, 28)]
COMBOS = ["PPPPNN", "NNNNDD", "DDDDDD", "NNNNNN", "PPPPDD", "DDNDDD", "NDDNDN", "DNNNDD"]


def rel_maxdiff(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", COMBOS)
def test_mixed_parity(bc, shape):
    print(f"mixed {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=sum(map(ord, bc)) + shape[0])
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, zero_mode=zero_mode)
        got = pa.solve(f, bc, zero_mode=zero_mode, precision=1)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} {shape} zero_mode={zero_mode}: rel max diff {err:.3e}"


# Short x lines (T lanes per line covering less than a 128-byte line) store through their LDS
# slices as one contiguous run per block, and inverse transforms load that way: lengths of
# every thread count that takes that path (FP64 up to T = 8 threads per line, FP32 up to 16) in
# every x kind, both precisions; 35 lines per grid, so each length's last block is ragged.
# Cells are cubes (the refinement's 7-point solution is then the spectral one to the bar).
STAGED_NX = [8, 12, 20, 28, 32, 48, 64, 80, 112, 128, 160, 224, 256, 320, 384, 448, 640]
X_KINDS = ["PPPPNN", "NNNNDD", "DDDDDD", "NDDNDN", "DNNNDD"]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nx", STAGED_NX)
@pytest.mark.parametrize("bc", X_KINDS)
def test_short_x_lines_staged(bc, nx, precision):
    shape = (7, 5, nx)
    print(f"short x lines {bc} {shape} precision={precision}", flush=True)
    f = prng.uniform_pm1(shape, seed=nx * 7 + precision + sum(map(ord, bc)))
    box = [0.0, 1.0, 0.0, 5.0 / nx, 0.0, 7.0 / nx]
    ref = oracle.solve(f, bc, box, zero_mode=1)  # the library's default (0,0)-mode rule
    got = pa.solve(f, bc, box, precision=precision)
    err = rel_maxdiff(got, ref)
    # refined mixed solves converge to the 7-point solution, ~1e-12 from the spectral one along
    # long periodic x axes (test_mixed_converges_to_stencil_solution)
    bar = 3e-12 if precision and bc[0] == "P" and nx >= 256 else TOL
    assert err <= bar, f"{bc} {shape} precision={precision}: rel max diff {err:.3e}"
    if precision:
        # the FP32 transforms alone (no refinement): a misplaced element would be O(1)
        got0 = pa.solve(f, bc, box, precision=1, refine=0)
        err0 = rel_maxdiff(got0, ref)
        assert err0 <= 1e-5, f"{bc} {shape} FP32 solve: rel max diff {err0:.3e}"


# The correction step fused into one kernel (the residual computed inside the FP32 x forward,
# PP_OPT_FUSE_RESIDUAL, default on where ny is a multiple of the kernel's lines per block):
# every fusing line length (32 .. 1024 points, lines per block 64 .. 4), ny = 8 x lines per block
# (the XCD-grouped block numbering) and fewer, every x kind, Dirichlet faces, the pinned
# column (its plane-0 sums come from the fused blocks) and periodic z, one to three steps (the
# first reads u0, later ones u0 + the correction sum).  Bar: the unfused path within 1e-13, and
# the oracle at 1e-12 after two or more steps, 1e-11 after one (or the unfused path's own error
# where that is larger).
FUSED_SHAPES = [(6, 64, 32), (5, 32, 128), (4, 128, 256), (4, 16, 256), (4, 64, 512), (5, 8, 512), (4, 4, 1024),
                (5, 32, 192), (4, 16, 640), (6, 16, 384), (4, 64, 96)]
FUSED_BCS = ["DDNDDD", "PPPPNN", "NNNNNN", "DDDDDD", "NDDNDN", "DNNNDD", "PPPPPP", "NNNNPP", "DDPPDD"]


@pytest.mark.parametrize("shape", FUSED_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", FUSED_BCS)
def test_fused_residual_step(bc, shape):
    nz, ny, nx = shape
    print(f"fused residual step {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=nx + 3 * ny + 7 * nz + sum(map(ord, bc)))
    box = [0.0, 1.0, 0.0, ny / nx, 0.0, nz / nx]
    ref = oracle.solve(f, bc, box, zero_mode=1)  # the library's default (0,0)-mode rule
    for refine in (1, 2, 3):
        fused = pa.solve(f, bc, box, precision=1, refine=refine)
        plain = pa.solve(f, bc, box, precision=1, refine=refine, fuse=0)
        d = rel_maxdiff(fused, plain)
        assert d <= 1e-13, f"{bc} {shape} refine={refine}: fused vs unfused {d:.3e}"
        err = rel_maxdiff(fused, ref)
        # the FP64 bar after two or more steps, 1e-11 after one -- or, where the unfused step
        # itself is further from the oracle, no worse than it: one FP32 correction on 512-wide
        # pinned grids (~2e-11), and after two the 7-point solution on a 1024-long periodic axis
        # (~1.7e-12 from the spectral one, as test_mixed_converges_to_stencil_solution explains)
        bar = max(TOL if refine >= 2 else 1e-11, 1.01 * rel_maxdiff(plain, ref))
        assert err <= bar, f"{bc} {shape} refine={refine}: fused vs oracle {err:.3e}"


def test_fused_residual_option():
    """PP_OPT_FUSE_RESIDUAL takes 0 or 1 only; the fused step's launch count (one kernel per
    correction instead of two) shows in the handle's launch marks"""
    f = prng.uniform_pm1((4, 16, 256), seed=5)
    with pa.PoissonHIP(256, 16, 4) as M:
        with pytest.raises(pa.PittPackError):
            M.set_option(pa.OPT_FUSE_RESIDUAL, 2)
        M.assignBoundary("DDNDDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        M.set_option(pa.OPT_PRECISION, 1)
        M.set_option(pa.OPT_TIMING, 1)
        counts = []
        for fuse in (1, 0):
            M.set_option(pa.OPT_FUSE_RESIDUAL, fuse)
            M.assignRhs(f)  # the solve consumes its right-hand side
            M.pittPack()
            counts.append(len(M.kernel_times()))
        assert counts[1] - counts[0] == 2, counts  # two refinement steps, one kernel fewer each


# periodic z: the z stage computes in FP64 on the FP32 work field (one-pass circulant solve for
# lengths with a compiled complex plan, real z-line transforms around a division otherwise);
# the singular column (x and y in {P, N}) follows the reference's averaged row 0
# (shermanMorrisonThomasV1) from FP64 plane sums, in the initial solve and every correction
ZPER_MIXED = ["PPPPPP", "NNNNPP", "NNPPPP", "DDPPPP", "DDDDPP", "DDNNPP", "NDDNPP", "DNNNPP"]


@pytest.mark.parametrize("nz", [4, 5, 8, 22, 40, 64, 101, 128, 512, 600])
@pytest.mark.parametrize("bc", ZPER_MIXED)
def test_mixed_periodic_z(bc, nz):
    shape = (nz, 12, 20)
    print(f"mixed periodic z {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=nz * 3 + sum(map(ord, bc)))
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    ref = oracle.solve(f, bc, box)
    got = pa.solve(f, bc, box, precision=1)
    err = rel_maxdiff(got, ref)
    assert err <= TOL, f"{bc} nz={nz}: rel max diff {err:.3e}"


@pytest.mark.parametrize("bc", ["PPPPPP", "NNNNPP"])
def test_mixed_periodic_z_contracts(bc):
    """periodic z on a larger grid: FP32 accuracy without refinement, the FP64 bar after two
    steps"""
    shape = (96, 64, 128)
    print(f"mixed periodic z refine {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=41)
    ref = oracle.solve(f, bc)
    errs = [rel_maxdiff(pa.solve(f, bc, precision=1, refine=r), ref) for r in (0, 1, 2)]
    assert errs[0] <= 1e-5, errs
    assert errs[1] <= 1e-10, errs
    assert errs[2] <= TOL, errs


@pytest.mark.parametrize("bc", ["PPPPPP", "NNNNPP"])
def test_mixed_periodic_z_thin_cells(bc):
    """cells six times thinner in z than in x and y make the low (kx, ky) columns
    ill-conditioned (|eig + 2| ~ 1e-5): the refinement converges to the exact 7-point solution
    (residual at rounding level, below the oracle's own), while the reference's spectral FP64
    solve, and so the oracle, carries cancellation error of about eps / |eig + 2| there"""
    shape = (96, 64, 128)
    box = [0.0, 2.0, 0.0, 1.0, 0.0, 0.25]
    print(f"mixed periodic z thin cells {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=41)
    fe = oracle.effective_rhs(f, bc, box)
    ref = oracle.solve(f, bc, box)
    got = pa.solve(f, bc, box, precision=1)

    def resid(u):
        r = fe - oracle.apply_operator(u, bc, box)
        r[0] -= r[0].mean()  # the gauge row's plane-0 defect (row 0 replaced by the averaged rule)
        return float(np.abs(r).max() / np.abs(fe).max())

    assert resid(got) <= 2e-13 and resid(got) <= resid(ref), (resid(got), resid(ref))
    assert rel_maxdiff(got, ref) <= 1e-11


# residual tiles at the grid's edge: the 64-column kernel where its last column or row is a
# tile's first (nx = 1 mod 64 is odd; ny = 1 mod 4) and the two-cells-per-lane kernel (taken
# where its 128-column tiles keep as many lanes busy: nx = 66, 128, 256) with ny = 1 mod 4
@pytest.mark.parametrize("shape", [(12, 9, 130), (7, 5, 258), (20, 13, 66), (6, 13, 128), (8, 12, 256), (9, 17, 65)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", ["PPPPNN", "DDDDDD", "NNNNDD", "DDNDDD", "NNPPPP"])
def test_mixed_residual_tile_edges(bc, shape):
    print(f"mixed residual tile edges {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=sum(map(ord, bc)) + shape[2])
    ref = oracle.solve(f, bc, zero_mode=1)
    got = pa.solve(f, bc, zero_mode=1, precision=1)
    assert rel_maxdiff(got, ref) <= TOL


@pytest.mark.parametrize("bc", ["PPPPNN", "NNNNDD", "DDNDDD"])
def test_mixed_refinement_contracts(bc):
    """FP32 accuracy without refinement, FP64 accuracy after one or two steps"""
    print(f"mixed refine {bc}", flush=True)
    f = prng.uniform_pm1((32, 48, 64), seed=5)
    ref = oracle.solve(f, bc, zero_mode=1)
    errs = [rel_maxdiff(pa.solve(f, bc, precision=1, refine=r), ref) for r in (0, 1, 2)]
    assert 1e-9 < errs[0] <= 1e-5, errs
    assert errs[1] <= 1e-11, errs
    assert errs[2] <= TOL, errs


@pytest.mark.parametrize("tau", [0.0, 2e-3, 0.5])
def test_mixed_tau(tau):
    """the FP64 column threshold changes the speed of refinement, not its limit"""
    print(f"mixed tau {tau}", flush=True)
    f = prng.uniform_pm1((24, 32, 32), seed=9)
    ref = oracle.solve(f, "NNNNDD", zero_mode=1)
    got = pa.solve(f, "NNNNDD", precision=1, refine=3, tau=tau)
    assert rel_maxdiff(got, ref) <= TOL


@pytest.mark.parametrize("nz", [600, 1000, 1024])
def test_mixed_long_columns(nz):
    """columns of 600 to 1024 rows (FP32 z solve with up to 64 rows per lane)"""
    print(f"mixed nz={nz}", flush=True)
    shape = (nz, 16, 16)
    box = [0.0, 1.0, 0.0, 1.0, 0.0, nz / 16.0]
    f = prng.uniform_pm1(shape, seed=nz)
    for bc in ("NNNNDD", "PPPPNN", "DDDDDN"):
        ref = oracle.solve(f, bc, box, zero_mode=1)
        got = pa.solve(f, bc, box, zero_mode=1, precision=1)
        err = rel_maxdiff(got, ref)
        assert err <= TOL, f"{bc} nz={nz}: rel max diff {err:.3e}"


def test_mixed_device_forms():
    """in-place and out-of-place device calls on torch buffers, stage timing record"""
    import torch

    print("mixed device forms", flush=True)
    dev = torch.device("cuda:0")
    shape = (32, 64, 64)
    f = prng.uniform_pm1(shape, seed=3)
    ref = oracle.solve(f, "DDNDDD", zero_mode=1)
    with pa.PoissonHIP(64, 64, 32) as M:
        M.assignBoundary("DDNDDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        M.set_option(pa.OPT_PRECISION, 1)
        M.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        M.set_option(pa.OPT_TIMING, 2)
        fd = torch.from_numpy(f).to(dev)
        ud = torch.empty_like(fd)
        M.solve_device_oop(fd.data_ptr(), ud.data_ptr())
        torch.cuda.synchronize()
        assert rel_maxdiff(ud.cpu().numpy(), ref) <= TOL
        assert np.array_equal(fd.cpu().numpy(), f)  # right-hand side kept
        kt = M.kernel_times()
        assert len(kt) >= 6 and all(t > 0 for t in kt)
        M.solve_device(fd.data_ptr())
        torch.cuda.synchronize()
        assert rel_maxdiff(fd.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_handle_graph_replays_match_eager(precision):
    """PP_OPT_GRAPH on a single-solve handle (the flow-solver pattern): repeated device solves
    on the same pointers replay a captured graph.  Every replay, with torch work (new
    right-hand sides) and synchronisations between them, must equal an eager handle's result
    bit for bit."""
    import torch

    print(f"handle graph replays precision={precision}", flush=True)
    dev = torch.device("cuda:0")
    nz, ny, nx = 40, 48, 64
    stream = torch.cuda.Stream(device=dev)
    handles = []
    for graph in (0, 1):
        M = pa.PoissonHIP(nx, ny, nz)
        M.assignBoundary("DDNDDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        if precision:
            M.set_option(pa.OPT_PRECISION, 1)
        M.set_option(pa.OPT_GRAPH, graph)
        M.set_stream(stream.cuda_stream)
        handles.append(M)
    E, G = handles
    try:
        with torch.cuda.stream(stream):
            f = torch.empty((nz, ny, nx), dtype=torch.float64, device=dev)
            ue = torch.empty_like(f)
            ug = torch.empty_like(f)
            for rep, seed in enumerate((1, 2, 1, 3, 2)):
                gen = torch.Generator(device=dev).manual_seed(seed)
                f.uniform_(-1, 1, generator=gen)
                E.solve_device_oop(f.data_ptr(), ue.data_ptr())
                G.solve_device_oop(f.data_ptr(), ug.data_ptr())
                stream.synchronize()
                assert torch.equal(ue, ug), f"solve {rep} (seed {seed}): graph replay != eager"
    finally:
        E.close()
        G.close()


def _stencil_solution(f, bc, box):
    """the 7-point ghost-cell system solved directly (sparse LU plus one refinement step with an
    extended-precision residual), independent of any spectral eigenvalue formula"""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    nz, ny, nx = f.shape
    d = [(box[1] - box[0]) / nx, (box[3] - box[2]) / ny, (box[5] - box[4]) / nz]

    def op(n, lo, hi):
        main = -2.0 * np.ones(n)
        main[0] += -1.0 if lo == "D" else 1.0
        main[-1] += -1.0 if hi == "D" else 1.0
        return sp.diags([np.ones(n - 1), main, np.ones(n - 1)], [-1, 0, 1], format="csr")

    Ax, Ay, Az = op(nx, bc[0], bc[1]), op(ny, bc[2], bc[3]), op(nz, bc[4], bc[5])
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    L = ((d[2] / d[0]) ** 2 * sp.kron(sp.kron(Iz, Iy), Ax) + (d[2] / d[1]) ** 2 * sp.kron(sp.kron(Iz, Ay), Ix)
         + sp.kron(sp.kron(Az, Iy), Ix)).tocsc()
    fe = oracle.effective_rhs(f, bc, box).ravel()
    lu = spl.splu(L)
    u = lu.solve(fe)
    r = fe.astype(np.longdouble) - L.astype(np.longdouble) @ u.astype(np.longdouble)
    return (u + lu.solve(r.astype(np.float64))).reshape(f.shape)


@pytest.mark.parametrize("bc", ["DDDDDD", "DDNNDD", "NNDDDD"])
def test_mixed_converges_to_stencil_solution(bc):
    """On cells much thinner than dz (here dy = dz / 250, so Cy = 62500) the reference's
    eigenvalues -2 + 2 cos(theta) lose digits to cancellation at small theta, which the FP64
    spectral solve (oracle and GPU alike) carries: it sits ~1e-12 off the exact solution of
    the 7-point system.  The mixed form's FP64-residual refinement converges to that exact
    solution instead.  Both facts are pinned against a direct sparse solve."""
    shape, box = (4, 1000, 24), [0, 1, 0, 1, 0, 1]
    print(f"mixed vs stencil {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=1003)
    exact = _stencil_solution(f, bc, box)
    ref = oracle.solve(f, bc, box, zero_mode=1)
    scale = np.abs(exact).max()
    u64 = pa.solve(f, bc, box, zero_mode=1)
    um = pa.solve(f, bc, box, zero_mode=1, precision=1)
    assert np.abs(u64 - ref).max() <= 1e-14 * scale         # FP64 form: the reference's arithmetic
    assert np.abs(ref - exact).max() >= 3e-13 * scale        # ... which is off the exact solution
    assert np.abs(um - exact).max() <= 1e-13 * scale         # mixed form: the exact solution


ZT_SHAPES = [(16, 64, 48), (40, 128, 96), (33, 48, 40), (200, 96, 64), (24, 512, 32), (700, 32, 48), (20, 1024, 32)]


@pytest.mark.parametrize("shape", ZT_SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("bc", ["PPPPNN", "DDNDDD", "NNNNDD", "NDDNNN"])
def test_mixed_ztile_layouts(bc, shape):
    """the mixed form's FP32 z tiles (32-column y blocks, both y shapes) and the work field in
    place: each meets the bar against the oracle"""
    print(f"mixed z layouts {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=len(bc) + shape[1])
    nz, ny, nx = shape
    # cubic cells: the refinement converges to the exact 7-point solution, which on cells much
    # thinner than dz differs from the spectral solution by its eigenvalue rounding
    # (test_mixed_converges_to_stencil_solution); with Neumann z faces the lowest modes of a
    # long axis carry the bound 4 eps (n / pi)^2 of that rounding
    h = 1.0 / max(shape)
    box = [0, nx * h, 0, ny * h, 0, nz * h]
    tol = TOL if "D" in bc[4:] else max(TOL, 4 * np.finfo(float).eps * (max(shape) / np.pi) ** 2)
    for zero_mode in (0, 1):
        ref = oracle.solve(f, bc, box, zero_mode=zero_mode)
        for zlayout, var in ((1, 0), (1, 1), (0, 0)):
            got = pa.solve(f, bc, box, zero_mode=zero_mode, precision=1, zlayout=zlayout, line_variant=var)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= tol, f"{bc} {shape} zm={zero_mode} zlayout={zlayout} variant={var}: {err:.3e}"


def test_mixed_autotune():
    """the mixed-form tuner picks a z layout / y shape and the result meets the bar"""
    shape = (64, 128, 96)
    f = prng.uniform_pm1(shape, seed=77)
    ref = oracle.solve(f, "DDNDDD", zero_mode=1)
    nz, ny, nx = shape
    with pa.PoissonHIP(nx, ny, nz) as M:
        M.set_option(pa.OPT_PRECISION, 1)
        M.set_option(pa.OPT_AUTOTUNE, 1)
        M.assignBoundary("DDNDDD")
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        M.assignRhs(f)
        M.pittPack()
        u = M.getSolution()
        t = M.tuning()
    print("mixed tuning", t, flush=True)
    assert t["mixed_ztiles"] in (0, 1) and t["mixed_y_variant"] in (0, 1)
    assert np.abs(u - ref).max() <= TOL * np.abs(ref).max()
