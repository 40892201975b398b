"""Batched form on the GPU: many independent FP64 solves of uneven sizes and boundary
conditions in one set of launches per stage and shape class.

Each member must be bitwise the single-solve handle's result for the same input (the batch
runs the same kernels with the same arithmetic), and within the FP64 bar (1e-12 relative
max-norm) of the oracle.  The device right-hand-side generator must reproduce the Python
counter generator bit for bit.
"""
import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (nx, ny, nz, bc, box): mixed line lengths (radix 2 and 3), ragged z, all x/y combos
MEMBERS = [
    (32, 48, 17, "PPPPNN", [0, 1, 0, 1, 0, 1]),
    (64, 32, 40, "NNNNDD", [0, 2, 0, 1, 0, 1]),
    (48, 48, 33, "DDDDDD", [0, 1, 0, 1, 0, 1]),
    (96, 64, 16, "DDNNNN", [0, 1, 0, 1, 0, 1]),
    (32, 32, 100, "PPPPDD", [0, 1, 0, 1, 0, 3]),
    (64, 96, 64, "DDNDDD", [0, 1, 0, 1, 0, 1]),
    (48, 32, 5, "NNNNNN", [0, 1, 0, 1, 0, 1]),
    (32, 64, 128, "NNDDND", [-1, 1, 0, 1, 0, 1]),
    (64, 32, 40, "PPPPNN", [0, 1, 0, 1, 0, 1]),
    (32, 48, 17, "DDDDNN", [0, 1, 0, 1, 0, 1]),
]


def rel_maxdiff(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_batch_matches_single_and_oracle():
    print("batch members", flush=True)
    with pa.Batch() as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in MEMBERS]
        B.finalize()
        assert B.count() == (len(MEMBERS), sum(nx * ny * nz for nx, ny, nz, _, _ in MEMBERS))
        x, y, z = B.launch_count()
        assert x == len({(m[0], m[3][:2]) for m in MEMBERS})
        assert y == len({(m[1], m[3][2:4]) for m in MEMBERS})
        assert z == len({(m[2] + 15) // 16 for m in MEMBERS})
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=100 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            got = B.get_solution(i)
            single = pa.solve(rhs[i], bc, box)
            assert np.array_equal(got, single), f"member {i} {bc}: batched != single ({rel_maxdiff(got, single):.2e})"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"member {i} {bc} {(nx, ny, nz)}: {err:.3e}"


# radix-5/7 and short lines (N = 8 .. 640) next to radix-2/3 ones in one batch
LENGTH_MEMBERS = [
    (16, 12, 10, "NNNNDD", [0, 1, 0, 1, 0, 1]),
    (24, 20, 18, "DDNDDD", [0, 1, 0, 1, 0, 1]),
    (40, 80, 12, "PPPPNN", [0, 2, 0, 1, 0, 1]),
    (640, 8, 4, "DDDDDD", [0, 1, 0, 1, 0, 1]),
    (160, 320, 4, "NDDNDN", [0, 1, 0, 1, 0, 1]),
    (20, 40, 33, "DDNNNN", [0, 1, 0, 1, 0, 1]),
    (112, 56, 9, "NNNNDD", [0, 1, 0, 2, 0, 1]),
    (32, 48, 17, "PPPPNN", [0, 1, 0, 1, 0, 1]),
]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_line_lengths(precision):
    print(f"batch line lengths precision={precision}", flush=True)
    with pa.Batch(precision=precision) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in LENGTH_MEMBERS]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, LENGTH_MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=300 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, LENGTH_MEMBERS):
            got = B.get_solution(i)
            if precision == 0:
                single = pa.solve(rhs[i], bc, box)
                assert np.array_equal(got, single), f"member {i}: batched != single"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"member {i} {bc} ({nx},{ny},{nz}): rel max diff {err:.3e}"


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_1024_point_y_lines(precision):
    """y lines of 1024 points: 16 FP64 columns per y block (one 1024-thread block), so FP64
    members take z tiles as for shorter lines"""
    print(f"batch 1024-point y lines precision={precision}", flush=True)
    members = [(16, 1024, 9, "DDNDDD", [0, 1, 0, 4, 0, 1]), (32, 64, 17, "PPPPNN", [0, 1, 0, 1, 0, 1])]
    with pa.Batch(precision=precision) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in members]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, members):
            f = prng.uniform_pm1((nz, ny, nx), seed=700 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, members):
            got = B.get_solution(i)
            if precision == 0:
                assert np.array_equal(got, pa.solve(rhs[i], bc, box)), f"member {i}: batched != single"
            err = rel_maxdiff(got, oracle.solve(rhs[i], bc, box, zero_mode=1))
            assert err <= TOL, f"member {i} {bc} ({nx},{ny},{nz}): rel max diff {err:.3e}"


# periodic z next to Dirichlet / Neumann z in one batch: lengths with a compiled complex plan
# (one launch per length over all such members), other lengths (real z-line transforms per
# member), the singular column (x and y in {P, N}); FP64 columns of 513..1024 rows
ZPER_MEMBERS = [
    (32, 48, 64, "PPPPPP", [0, 1, 0, 1, 0, 1]),
    (20, 12, 22, "NNNNPP", [0, 1, 0, 1, 0, 2]),
    (48, 32, 64, "DDDDPP", [0, 2, 0, 1, 0, 1]),
    (24, 20, 101, "NNPPPP", [0, 1, 0, 1, 0, 3]),
    (33, 16, 8, "NDDNPP", [0, 1, 0, 1, 0, 1]),
    (16, 24, 128, "DDPPPP", [0, 1, 0, 1, 0, 4]),
    (32, 32, 40, "PPPPNN", [0, 1, 0, 1, 0, 1]),
    (16, 16, 600, "NNNNDD", [0, 1, 0, 1, 0, 6]),
    (16, 32, 1024, "DDNDDD", [0, 1, 0, 1, 0, 8]),
    (64, 48, 17, "DDDDNN", [0, 1, 0, 1, 0, 1]),
]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_batch_periodic_z_and_long_columns(precision):
    print(f"batch periodic z / long columns precision={precision}", flush=True)
    with pa.Batch(precision=precision) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in ZPER_MEMBERS]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, ZPER_MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=300 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, ZPER_MEMBERS):
            got = B.get_solution(i)
            if precision == 0:
                single = pa.solve(rhs[i], bc, box)
                assert np.array_equal(got, single), f"member {i} {bc}: batched != single ({rel_maxdiff(got, single):.2e})"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"member {i} {bc} {(nx, ny, nz)}: {err:.3e}"


@pytest.mark.parametrize("zthomas", [0.0, 0.5])
def test_batch_zthomas(zthomas):
    """the FP64 batch's nearly singular columns by the reference's Thomas sweeps (threshold 0.5:
    many columns per member) or none by them (0): each member bitwise the single handle with
    the same threshold, and within the FP64 bar of the oracle"""
    print(f"batch zthomas={zthomas}", flush=True)
    with pa.Batch(zthomas=zthomas) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in MEMBERS]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=700 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            got = B.get_solution(i)
            with pa.PoissonHIP(nx, ny, nz) as M:
                M.set_option_double(pa.OPT_ZTHOMAS_TAU, zthomas)
                M.assignBoundary(bc)
                M.setBox(box)
                M.setCoords(2)
                M.assignRhs(rhs[i])
                M.pittPack()
                single = M.getSolution()
            assert np.array_equal(got, single), f"member {i} {bc}: batched != single ({rel_maxdiff(got, single):.2e})"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            assert rel_maxdiff(got, ref) <= TOL, f"member {i} {bc}"


def test_batch_repeated_and_device_fill():
    """device generator == Python generator; the right-hand sides are kept, so a second solve
    reproduces the first"""
    print("batch fill", flush=True)
    with pa.Batch() as B:
        for nx, ny, nz, bc, box in MEMBERS[:4]:
            B.add(nx, ny, nz, bc, box)
        B.finalize()
        B.fill_uniform(7)
        for i in range(4):
            nx, ny, nz = MEMBERS[i][:3]
            assert np.array_equal(B.get_rhs(i), prng.uniform_pm1((nz, ny, nx), seed=7 + i))
        B.solve()
        first = [B.get_solution(i) for i in range(4)]
        B.set_option(pa.OPT_TIMING, 1)
        B.solve()
        for i in range(4):
            assert np.array_equal(B.get_solution(i), first[i])
        kt = B.kernel_times()
        assert len(kt) == 5 and all(t > 0 for t in kt)


def test_batch_errors():
    print("batch errors", flush=True)
    with pa.Batch() as B:
        with pytest.raises(pa.PittPackError) as e:
            B.solve()
        assert e.value.code == pa.CALL_ORDER
        with pytest.raises(pa.PittPackError) as e:
            B.add(1030, 32, 16, "PPPPNN")
        assert e.value.code == pa.UNSUPPORTED_SIZE
        with pytest.raises(pa.PittPackError) as e:
            B.add(32, 32, 16, "PNPPNN")
        assert e.value.code == pa.UNSUPPORTED_BC
        B.add(32, 32, 16, "PPPPNN")
        B.finalize()
        with pytest.raises(pa.PittPackError) as e:
            B.add(32, 32, 16, "PPPPNN")
        assert e.value.code == pa.CALL_ORDER


@pytest.mark.parametrize("zero_mode", [0, 1])
def test_batch_mixed_precision(zero_mode):
    """mixed form over a batch: FP64 bar against the oracle after the default refinement,
    and equal to the single-solve mixed handle"""
    print(f"batch mixed zero_mode={zero_mode}", flush=True)
    with pa.Batch(precision=1, zero_mode=zero_mode) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in MEMBERS]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            f = prng.uniform_pm1((nz, ny, nx), seed=300 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.set_option(pa.OPT_TIMING, 1)
        B.solve()
        assert len(B.kernel_times()) == 5 + 6 * 2
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
            got = B.get_solution(i)
            ref = oracle.solve(rhs[i], bc, box, zero_mode=zero_mode)
            err = rel_maxdiff(got, ref)
            assert err <= TOL, f"member {i} {bc} {(nx, ny, nz)}: {err:.3e}"
            single = pa.solve(rhs[i], bc, box, zero_mode=zero_mode, precision=1)
            assert rel_maxdiff(got, single) <= 1e-14, f"member {i}: batched vs single mixed"


@pytest.mark.parametrize("refine", [1, 3])
def test_batch_mixed_refine_counts(refine):
    """the solution's FP32 parts with one refinement step (u0 + c, written in FP64 at once) and
    with three (the second correction summed into the first in FP32): each member agrees with
    the single-solve mixed handle (whose FP32 y stages may take 32-column z tiles: the same
    answer to the refinement's accuracy) and with the oracle, at the FP64 bar after three steps
    and at FP32 squared (1e-11) after one; the refinement count may change after finalize"""
    print(f"batch mixed refine={refine}", flush=True)
    with pa.Batch(precision=1) as B:
        ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in MEMBERS[:6]]
        B.finalize()
        rhs = []
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS[:6]):
            f = prng.uniform_pm1((nz, ny, nx), seed=500 + i)
            rhs.append(f)
            B.set_rhs(i, f)
        B.solve()
        B.set_option(pa.OPT_REFINE, refine)
        B.solve()
        for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS[:6]):
            got = B.get_solution(i)
            bar = TOL if refine >= 2 else 1e-11
            single = pa.solve(rhs[i], bc, box, precision=1, refine=refine)
            assert rel_maxdiff(got, single) <= bar, f"member {i}: batched vs single {rel_maxdiff(got, single):.2e}"
            ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
            assert rel_maxdiff(got, ref) <= bar, f"member {i} {bc}: {rel_maxdiff(got, ref):.2e}"


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_graph_replays_match_eager(precision):
    """PP_OPT_GRAPH: the second solve captures the batch's launches into a HIP graph and later
    solves replay it.  Every replay, with eager work (new right-hand sides from the device
    generator, copies and synchronisations) between them, must equal the eager batch bit for
    bit."""
    print(f"batch graph replays precision={precision}", flush=True)
    E = pa.Batch(precision=precision, graph=0)
    G = pa.Batch(precision=precision, graph=1)
    try:
        for B in (E, G):
            for nx, ny, nz, bc, box in MEMBERS:
                B.add(nx, ny, nz, bc, box)
            B.finalize()
        for rep, seed in enumerate((11, 12, 11, 13, 12)):
            for B in (E, G):
                B.fill_uniform(seed, round_f32=bool(precision))
                B.solve()
                B.synchronize()
            for i in range(len(MEMBERS)):
                a, b = G.get_solution(i), E.get_solution(i)
                assert np.array_equal(a, b), f"solve {rep} (seed {seed}) member {i}: graph replay != eager"
    finally:
        E.close()
        G.close()


def _batch_vs_oracle(B, tag):
    ids = [B.add(nx, ny, nz, bc, box) for nx, ny, nz, bc, box in MEMBERS]
    B.finalize()
    rhs = []
    for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
        f = prng.uniform_pm1((nz, ny, nx), seed=300 + i)
        rhs.append(f)
        B.set_rhs(i, f)
    B.solve()
    for i, (nx, ny, nz, bc, box) in zip(ids, MEMBERS):
        ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
        err = rel_maxdiff(B.get_solution(i), ref)
        assert err <= TOL, f"{tag} member {i} {bc} ({nx},{ny},{nz}): rel max diff {err:.3e}"


def test_batch_alternative_shapes_vs_oracle():
    """every FP64 line class of the batch in its alternative launch shape (x classes then use
    their own block items)"""
    print("batch alternative shapes", flush=True)
    with pa.Batch(line_variant=1) as B:
        _batch_vs_oracle(B, "alternative shapes")


def test_batch_autotune_vs_oracle():
    """PP_OPT_AUTOTUNE on a batch: each line class and direction timed in both shapes at
    finalize, the faster kept; the tuned batch still meets the FP64 bar"""
    print("batch autotune", flush=True)
    with pa.Batch(autotune=1) as B:
        _batch_vs_oracle(B, "autotuned")
        n = B.tuning()
        assert 0 <= n <= 2 * (B.launch_count()[0] + B.launch_count()[1])


# mixed-form batch members whose corrections are fused into the x forward (ny a multiple of the
# fused kernel's lines per block: 64 for 32-point lines, 32 for 128, 16 for 256, 8 for 512)
# next to members that are not (ny = 48, 20, 12): same classes split into fused and plain parts
FUSED_MEMBERS = [
    (32, 64, 6, "DDNDDD", [0.0, 1.0, 0.0, 2.0, 0.0, 0.1875]),
    (32, 48, 5, "DDNDDD", [0.0, 1.0, 0.0, 1.5, 0.0, 0.15625]),
    (128, 32, 5, "PPPPNN", [0.0, 1.0, 0.0, 0.25, 0.0, 0.0390625]),
    (128, 20, 4, "PPPPNN", [0.0, 1.0, 0.0, 0.15625, 0.0, 0.03125]),
    (256, 16, 4, "NNNNNN", [0.0, 1.0, 0.0, 0.0625, 0.0, 0.015625]),
    (256, 12, 4, "NDDNDN", [0.0, 1.0, 0.0, 0.046875, 0.0, 0.015625]),
    (512, 8, 4, "DDDDDD", [0.0, 1.0, 0.0, 0.015625, 0.0, 0.0078125]),
    (192, 32, 5, "NNNNPP", [0.0, 1.0, 0.0, 1.0 / 6, 0.0, 5.0 / 192]),
]


def test_batch_fused_corrections():
    """the batched mixed form with its corrections fused per member (the single-solve handle's
    rule): each member within 1e-14 of the single handle and at the FP64 bar against the oracle
    (zero_mode 1, the library's default), and within 1e-13 of the same batch with the fusion
    off (PP_OPT_FUSE_RESIDUAL = 0 before finalize); a graph replay equals the eager solve"""
    print("batch fused corrections", flush=True)
    rhs = [prng.uniform_pm1((nz, ny, nx), seed=900 + i) for i, (nx, ny, nz, _, _) in enumerate(FUSED_MEMBERS)]
    out = {}
    for fuse in (1, 0):
        with pa.Batch(precision=1, fuse=fuse) as B:
            for nx, ny, nz, bc, box in FUSED_MEMBERS:
                B.add(nx, ny, nz, bc, box)
            B.finalize()
            for i, f in enumerate(rhs):
                B.set_rhs(i, f)
            B.solve()
            first = [B.get_solution(i) for i in range(len(rhs))]
            B.solve()  # captured into the graph
            B.solve()  # replayed
            for i in range(len(rhs)):
                assert np.array_equal(B.get_solution(i), first[i]), f"fuse={fuse} member {i}: replay != eager"
            out[fuse] = first
    for i, (nx, ny, nz, bc, box) in enumerate(FUSED_MEMBERS):
        got = out[1][i]
        single = pa.solve(rhs[i], bc, box, precision=1)
        assert rel_maxdiff(got, single) <= 1e-14, f"member {i}: batched vs single {rel_maxdiff(got, single):.2e}"
        plain = out[0][i]
        assert rel_maxdiff(got, plain) <= 1e-13, f"member {i}: fused vs unfused {rel_maxdiff(got, plain):.2e}"
        ref = oracle.solve(rhs[i], bc, box, zero_mode=1)
        bar = max(TOL, 1.01 * rel_maxdiff(plain, ref))
        assert rel_maxdiff(got, ref) <= bar, f"member {i} {bc}: {rel_maxdiff(got, ref):.2e}"
    with pa.Batch(precision=1) as B:
        B.add(32, 32, 4, "PPPPNN")
        B.finalize()
        with pytest.raises(pa.PittPackError):
            B.set_option(pa.OPT_FUSE_RESIDUAL, 0)  # after finalize
