"""pp_solve_device_oop_noalloc (no counterpart in the reference; for callers that cannot
allocate device memory) against pp_solve_device_oop, the entry point it is a variant of, on
torch device arrays through the C ABI.  Two calls in a row, the second on a smaller grid (the
tables in the shared scratch array are rebuilt), then the first shape again, periodic z,
arbitrary-length lines, a repeated call that reuses the resident tables, and the size and
argument errors.  Bar: bitwise equal to the handle on the same (in-place) z layout with the
PCR solve on every column (PP_OPT_ZTHOMAS_TAU = 0: this entry point does not take the Thomas
columns), and within 1e-12 of the handle's default configuration."""
import ctypes

import numpy as np
import pytest

import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu


def lib():
    return pa.load_library()


def noalloc(shape, bc, box, f, zero_mode=1):
    import torch

    nz, ny, nx = shape
    d_f = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    d_u = torch.full_like(d_f, float("nan"))
    X = np.ascontiguousarray(box, dtype=np.float64)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib().pp_solve_device_oop_noalloc(nx, ny, nz, bc.encode(), X.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           zero_mode, d_f.data_ptr(), d_u.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rc == 0, pa.error_name(rc)
    assert np.array_equal(d_f.cpu().numpy(), f)  # right-hand side kept
    return d_u.cpu().numpy()


def handle(shape, bc, box, f, zlayout, zero_mode=1, zthomas=None):
    import torch

    nz, ny, nx = shape
    d_f = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    d_u = torch.empty_like(d_f)
    with pa.PoissonHIP(nx, ny, nz) as M:
        M.set_option(pa.OPT_ZLAYOUT, zlayout)
        M.set_option(pa.OPT_ZERO_MODE, zero_mode)
        if zthomas is not None:
            M.set_option_double(pa.OPT_ZTHOMAS_TAU, zthomas)
        M.assignBoundary(bc)
        M.setBox(box)
        M.setCoords(2)
        torch.cuda.synchronize()
        M.solve_device_oop(d_f.data_ptr(), d_u.data_ptr())
        torch.cuda.synchronize()
    return d_u.cpu().numpy()


CASES = [
    ((40, 48, 64), "DDNDDD", [0, 1.5, 0, 1, -0.5, 0.5], 0),
    ((18, 20, 24), "PPPPNN", [0, 1, 0, 1, 0, 1], 0),  # smaller: the second call
    ((40, 48, 64), "DDNDDD", [0, 1.5, 0, 1, -0.5, 0.5], 0),  # back to the first
    ((24, 33, 37), "NDDNDD", [0, 1, 0, 1, 0, 1], 1),  # arbitrary-length lines
    ((22, 16, 12), "NNNNPP", [0, 1, 0, 1, 0, 1], 1),  # periodic z, real z-line transforms
    ((32, 16, 20), "PPPPPP", [0, 1, 0, 1, 0, 1], 1),  # periodic z, complex z FFT, singular column
    ((600, 8, 16), "NNNNNN", [0, 1, 0, 1, 0, 1], 0),  # long columns, pinned (0,0) column
]


def test_noalloc_matches_handle_in_sequence():
    prev = None
    for i, (shape, bc, box, zm) in enumerate(CASES):
        print(f"noalloc call {i}: {bc} {shape} zero_mode={zm}", flush=True)
        f = prng.uniform_pm1(shape, seed=40 + i)
        u = noalloc(shape, bc, box, f, zm)
        assert np.isfinite(u).all()
        ref_inplace = handle(shape, bc, box, f, 0, zm, zthomas=0.0)
        assert np.array_equal(u, ref_inplace), f"call {i}: differs from the in-place handle"
        ref = handle(shape, bc, box, f, 1, zm)
        assert np.abs(u - ref).max() <= 1e-12 * np.abs(ref).max(), f"call {i}"
        if prev is not None:
            assert prev != (shape, bc)
        prev = (shape, bc)


def test_noalloc_repeated_call_reuses_tables():
    shape, bc, box, zm = CASES[0]
    print("noalloc repeated", flush=True)
    f = prng.uniform_pm1(shape, seed=7)
    a = noalloc(shape, bc, box, f, zm)
    b = noalloc(shape, bc, box, f, zm)
    assert np.array_equal(a, b)


def test_noalloc_errors():
    import torch

    d = torch.zeros(8 * 8 * 8, dtype=torch.float64, device="cuda")
    X = np.array([0, 1, 0, 1, 0, 1], dtype=np.float64)
    Xp = X.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    s = torch.cuda.current_stream().cuda_stream
    L = lib()
    assert L.pp_solve_device_oop_noalloc(1025, 8, 8, b"NNNNDD", Xp, 1, d.data_ptr(), d.data_ptr(), s) == pa.UNSUPPORTED_SIZE
    assert L.pp_solve_device_oop_noalloc(8, 8, 2000, b"NNNNDD", Xp, 1, d.data_ptr(), d.data_ptr(), s) == pa.UNSUPPORTED_SIZE
    assert L.pp_solve_device_oop_noalloc(8, 8, 8, b"NNXXDD", Xp, 1, d.data_ptr(), d.data_ptr(), s) == pa.UNSUPPORTED_BC
    assert L.pp_solve_device_oop_noalloc(8, 8, 8, b"NNNNDD", Xp, 2, d.data_ptr(), d.data_ptr(), s) == pa.INPUT_ARGS
    assert L.pp_solve_device_oop_noalloc(8, 8, 8, b"NNNNDD", Xp, 1, None, d.data_ptr(), s) == pa.INPUT_ARGS
