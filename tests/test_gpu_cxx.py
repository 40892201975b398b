"""The C++ binding (include/pittpack_hip.hpp: class PoissonHIP and make_Poisson with the
reference's method names) driven in main.cpp's call order (src/main.cpp:73-118) from compiled
C++ (pittpack_amd/cxx/pp_cxx_check.cpp, built by g++), checked against the oracle and against
the Python binding's getError."""
import ctypes
import os

import numpy as np
import pytest

import oracle
import pittpack_amd as pa
from pittpack_amd import rng as prng

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(pa.LIB_PATH), "libpp_cxx_check.so")


def cxx():
    pa.load_library()  # the HIP library (and torch's runtime) first
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.pp_cxx_main_case.argtypes = [ctypes.c_int] * 3 + [ctypes.c_char_p, dp, dp, dp, dp]
    L.pp_cxx_main_case.restype = ctypes.c_int
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("bc,shape", [("NNNNDD", (24, 20, 16)), ("DDNDDD", (32, 48, 64)), ("PPPPNN", (16, 16, 33))])
def test_cxx_binding_user_rhs(bc, shape):
    nz, ny, nx = shape
    print(f"c++ binding {bc} {shape}", flush=True)
    f = prng.uniform_pm1(shape, seed=11)
    X = np.array([0, 1, 0, 1, 0, 1], dtype=np.float64)
    u = np.empty(shape)
    err = ctypes.c_double(-1)
    rc = cxx().pp_cxx_main_case(nx, ny, nz, bc.encode(), _dp(X), _dp(f), _dp(u), ctypes.byref(err))
    assert rc == 0, pa.error_name(rc)
    ref = oracle.solve(f, bc, zero_mode=1)
    assert np.abs(u - ref).max() <= 1e-12 * np.abs(ref).max()
    # getError of a user right-hand side: RMS against cos(2 pi z)
    z = (np.arange(nz) + 0.5) / nz
    want = np.sqrt(np.mean((u - np.cos(2 * np.pi * z)[:, None, None]) ** 2))
    assert abs(err.value - want) <= 1e-12 * want


def test_cxx_binding_manufactured_case():
    """rhs = NULL: initializeTrigonometric + getError, as the reference's default build runs"""
    nx = ny = nz = 32
    X = np.array([0, 1, 0, 1, 0, 1], dtype=np.float64)
    u = np.empty((nz, ny, nx))
    err = ctypes.c_double(-1)
    rc = cxx().pp_cxx_main_case(nx, ny, nz, b"NNNNDD", _dp(X), None, _dp(u), ctypes.byref(err))
    assert rc == 0, pa.error_name(rc)
    M = pa.PoissonHIP(nx, ny, nz)
    M.assignBoundary("NNNNDD")
    M.setBox(X)
    M.setCoords(2)
    M.initializeTrigonometric(0)
    M.pittPack()
    assert np.array_equal(M.getSolution(), u)
    assert M.getError() == err.value and 0 < err.value < 2e-3


def test_cxx_binding_errors_are_codes():
    X = np.array([0, 1, 0, 1, 0, 1], dtype=np.float64)
    u = np.empty((8, 8, 8))
    err = ctypes.c_double(-1)
    rc = cxx().pp_cxx_main_case(8, 8, 8, b"NNXXDD", _dp(X), _dp(u), _dp(u), ctypes.byref(err))
    assert rc == pa.UNSUPPORTED_BC
    rc = cxx().pp_cxx_main_case(1, 8, 8, b"NNNNDD", _dp(X), _dp(u), _dp(u), ctypes.byref(err))
    assert rc == pa.INPUT_ARGS
