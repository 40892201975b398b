"""CPU oracle for the PittPack solve path -- TEST INFRASTRUCTURE ONLY.

ctypes bindings for ``oracle/libpporacle.so`` (built from ``oracle/pp_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker: the product path
(``pittpack_amd``) never imports it.

Every function here restates a piece of the reference algorithm; the C file cites the
reference lines (src/pencilDcmp.cpp, src/signalProc.cpp, src/triDiag.cpp).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpporacle.so")
_lib = None

# line-transform kinds understood by ppo_line_transform
DST2, DCT2, DST3H, DCT3H, DCT4, DCT4H, DST4, DST4H = range(8)


def build(force: bool = False) -> str:
    """Compile the oracle with gcc via oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.ppo_fft.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.ppo_line_transform.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.ppo_zsolve_line.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_char, ctypes.c_char, ctypes.c_int, dp]
        L.ppo_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, dp, dp, ctypes.c_int]
        L.ppo_solve.restype = ctypes.c_int
        L.ppo_solve_ws.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, dp, dp, ctypes.c_int,
                                   dp]
        L.ppo_solve_ws.restype = ctypes.c_int
        L.ppo_set_per_line_plans.argtypes = [ctypes.c_int]
        L.ppo_effective_rhs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, dp, dp]
        L.ppo_effective_rhs.restype = ctypes.c_int
        L.ppo_apply_operator.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, dp, dp]
        L.ppo_apply_operator.restype = ctypes.c_int
        L.ppo_eigenvalue.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, dp, ctypes.c_int, ctypes.c_int]
        L.ppo_threads.argtypes = []
        L.ppo_threads.restype = ctypes.c_int
        L.ppo_eigenvalue.restype = ctypes.c_double
        _lib = L
    return _lib


def threads() -> int:
    """OpenMP threads the oracle's line, plane and column loops use."""
    return int(lib().ppo_threads())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised complex DFT of one line (sign -1 forward, +1 backward)."""
    z = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    buf = z.view(np.float64)
    lib().ppo_fft(len(z), _dp(buf), int(sign))
    return z


def line_transform(x: np.ndarray, kind: int) -> np.ndarray:
    """Apply one reference line transform (DST2/DCT2/... constants above) to a real line."""
    y = np.ascontiguousarray(np.asarray(x, dtype=np.float64)).copy()
    lib().ppo_line_transform(len(y), int(kind), _dp(y))
    return y


def zsolve_line(rhs: np.ndarray, eig: float, bcz: str, zero_mode: int = 0) -> np.ndarray:
    """One tridiagonal z solve as solveThmBatch does per (kx, ky) column."""
    y = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64)).copy()
    lib().ppo_zsolve_line(len(y), float(eig), bcz[0].encode(), bcz[1].encode(), int(zero_mode), _dp(y))
    return y


def _box(box):
    return np.ascontiguousarray(np.asarray(box if box is not None else [0, 1, 0, 1, 0, 1], dtype=np.float64))


def solve(rhs: np.ndarray, bc: str, box=None, zero_mode: int = 0, work: np.ndarray | None = None,
          out: np.ndarray | None = None) -> np.ndarray:
    """Whole FP64 solve.  rhs has shape (nz, ny, nx) (x fastest), unscaled.  work: optional
    workspace from `workspace(shape)` (repeated solves then allocate nothing); out: optional
    solution array."""
    f = np.ascontiguousarray(rhs, dtype=np.float64)
    nz, ny, nx = f.shape
    X = _box(box)
    u = out if out is not None else np.empty_like(f)
    if work is not None:
        if work.size < 2 * f.size or work.dtype != np.float64 or not work.flags.c_contiguous:
            raise ValueError("workspace too small")
        rc = lib().ppo_solve_ws(nx, ny, nz, bc.encode(), _dp(X), _dp(f), _dp(u), int(zero_mode), _dp(work))
    else:
        rc = lib().ppo_solve(nx, ny, nz, bc.encode(), _dp(X), _dp(f), _dp(u), int(zero_mode))
    if rc != 0:
        raise ValueError(f"oracle solve failed with code {rc} for bc={bc}")
    return u


def workspace(shape) -> np.ndarray:
    """Workspace of one solve of this (nz, ny, nx) shape: the reference's complex-interleaved
    field (2 doubles per point), touched once here."""
    w = np.empty(2 * int(np.prod(shape)), dtype=np.float64)
    w[:] = 0.0
    return w


def set_per_line_plans(on: bool) -> None:
    """Make and destroy the FFT plan of every line, as the reference's CPU path does with
    FFTW (src/poissonCPU.cpp:52-63); off (default): one plan per line set.  Same results."""
    lib().ppo_set_per_line_plans(int(bool(on)))


def effective_rhs(rhs: np.ndarray, bc: str, box=None) -> np.ndarray:
    f = np.ascontiguousarray(rhs, dtype=np.float64)
    nz, ny, nx = f.shape
    X = _box(box)
    out = np.empty_like(f)
    rc = lib().ppo_effective_rhs(nx, ny, nz, bc.encode(), _dp(X), _dp(f), _dp(out))
    if rc != 0:
        raise ValueError(f"oracle effective_rhs failed with code {rc}")
    return out


def apply_operator(u: np.ndarray, bc: str, box=None) -> np.ndarray:
    """dz^2-scaled 7-point operator with the ghost-cell closures the solve inverts."""
    v = np.ascontiguousarray(u, dtype=np.float64)
    nz, ny, nx = v.shape
    X = _box(box)
    out = np.empty_like(v)
    rc = lib().ppo_apply_operator(nx, ny, nz, bc.encode(), _dp(X), _dp(v), _dp(out))
    if rc != 0:
        raise ValueError(f"oracle apply_operator failed with code {rc}")
    return out


def eigenvalue(shape, bc: str, i: int, j: int, box=None) -> float:
    nz, ny, nx = shape
    X = _box(box)
    return lib().ppo_eigenvalue(nx, ny, nz, bc.encode(), _dp(X), int(i), int(j))
