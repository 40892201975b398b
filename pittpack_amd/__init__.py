"""pittpack_amd -- MI355X (gfx950) implementation of PittPack's solve path.

The compute lives in ``pittpack_amd/lib/libpittpack_hip.so`` (hand-written HIP kernels behind
the C ABI declared in ``include/pittpack_hip.h``).  This module is a thin ctypes binding plus a
Python mirror of the reference's ``PencilDcmp`` set-up/solve interface
(src/include/pencilDcmp.hpp:441-466), used by the tests, ``bench.py`` and Python callers.

There is no CPU fallback: if the shared library is missing or no GPU is visible, every solve
raises ``PittPackError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpittpack_hip.so")

SUCCESS = 0
ALLOCATION_FAIL = 15
INPUT_ARGS = 19
HIP_ERROR = 24
UNSUPPORTED_BC = 25
UNSUPPORTED_SIZE = 26
INVALID_HANDLE = 27
CALL_ORDER = 28
NO_DEVICE = 29
IO_FAIL = 30
NO_HDF5 = 31

IO_HDF5 = 0
IO_RAW = 1

OPT_ZERO_MODE = 1
OPT_TIMING = 2
OPT_ZSOLVER = 3
OPT_PRECISION = 4
OPT_REFINE = 5
OPT_TAU = 6
OPT_BATCH_MEMBERS = 7
OPT_ZLAYOUT = 8
OPT_MG_TOL = 9
OPT_GRAPH = 10
OPT_AUTOTUNE = 11
OPT_LINE_VARIANT = 12
OPT_MAX_OVERTAKE = 13
OPT_IO_FORMAT = 14
OPT_ZTHOMAS_TAU = 15
OPT_FUSE_RESIDUAL = 16

# every symbol include/pittpack_hip.h declares
EXPORTED = (
    "pp_create", "pp_destroy", "pp_assign_boundary", "pp_set_box", "pp_set_coords",
    "pp_assign_rhs", "pp_assign_rhs_device", "pp_pittpack", "pp_get_solution",
    "pp_get_solution_device", "pp_io", "pp_solve_device", "pp_solve_device_oop", "pp_set_stream", "pp_set_option",
    "pp_get_stage_times", "pp_get_device_field", "pp_get_error_enum", "pp_size_supported",
    "pp_selftest", "pp_set_option_double", "pp_get_kernel_times", "pp_get_tuning",
    "pp_batch_create", "pp_batch_destroy", "pp_batch_add", "pp_batch_finalize", "pp_batch_count",
    "pp_batch_launch_count", "pp_batch_field", "pp_batch_set_rhs", "pp_batch_get_solution",
    "pp_batch_fill_uniform", "pp_batch_set_stream", "pp_batch_set_option", "pp_batch_solve",
    "pp_batch_get_kernel_times", "pp_batch_synchronize", "pp_batch_solution", "pp_batch_set_option_double", "pp_batch_get_rhs",
    "pp_server_create", "pp_server_destroy", "pp_server_set_option", "pp_server_submit", "pp_server_wait",
    "pp_server_stats", "pp_batch_tuning", "pp_server_wait_ex", "pp_server_footprint", "pp_copy_device",
    "pp_init_trigonometric", "pp_get_error", "pp_fill_trigonometric", "pp_write_solution", "pp_hdf5_available",
    "pp_solve_device_oop_noalloc",
)


class PittPackError(RuntimeError):
    def __init__(self, code: int, where: str = ""):
        self.code = code
        name = error_name(code) if _lib is not None else str(code)
        super().__init__(f"{where}: PittPack error {code} ({name})")


_lib = None


def load_library(path: str | None = None):
    """Load the HIP library (no GPU needed to load it)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise PittPackError(HIP_ERROR, f"HIP library not built ({p}); run __graft_entry__.build()")
    # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7).  Whichever
    # is loaded first serves the whole process, and torch cannot initialise its GPU support
    # on the system runtime, so when torch is installed it is loaded before this library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(p)
    vp, i, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.pp_create.argtypes = [i, i, i, ctypes.POINTER(vp)]
    L.pp_destroy.argtypes = [vp]
    L.pp_assign_boundary.argtypes = [vp, ctypes.c_char_p]
    L.pp_set_box.argtypes = [vp, dp]
    L.pp_set_coords.argtypes = [vp, i]
    L.pp_assign_rhs.argtypes = [vp, dp]
    L.pp_assign_rhs_device.argtypes = [vp, vp]
    L.pp_pittpack.argtypes = [vp]
    L.pp_get_solution.argtypes = [vp, dp]
    L.pp_get_solution_device.argtypes = [vp, vp]
    L.pp_io.argtypes = [vp, ctypes.c_char_p, i]
    L.pp_solve_device.argtypes = [vp, vp]
    L.pp_solve_device_oop.argtypes = [vp, vp, vp]
    L.pp_set_stream.argtypes = [vp, vp]
    L.pp_set_option.argtypes = [vp, i, ctypes.c_longlong]
    L.pp_get_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.pp_set_option_double.argtypes = [vp, i, ctypes.c_double]
    L.pp_get_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_float), i, ctypes.POINTER(i)]
    L.pp_get_tuning.argtypes = [vp, ctypes.POINTER(i), i]
    L.pp_batch_tuning.argtypes = [vp, ctypes.POINTER(i)]
    L.pp_get_device_field.argtypes = [vp, ctypes.POINTER(vp)]
    L.pp_get_error_enum.argtypes = [i]
    L.pp_get_error_enum.restype = ctypes.c_char_p
    L.pp_size_supported.argtypes = [i]
    L.pp_selftest.argtypes = []
    ip, fp = ctypes.POINTER(i), ctypes.POINTER(ctypes.c_float)
    L.pp_batch_create.argtypes = [ctypes.POINTER(vp)]
    L.pp_batch_destroy.argtypes = [vp]
    L.pp_batch_add.argtypes = [vp, i, i, i, ctypes.c_char_p, dp, ip]
    L.pp_batch_finalize.argtypes = [vp]
    L.pp_batch_count.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_longlong)]
    L.pp_batch_launch_count.argtypes = [vp, ip]
    L.pp_batch_field.argtypes = [vp, i, ctypes.POINTER(vp)]
    L.pp_batch_set_rhs.argtypes = [vp, i, dp]
    L.pp_batch_get_solution.argtypes = [vp, i, dp]
    L.pp_batch_fill_uniform.argtypes = [vp, ctypes.c_ulonglong, i]
    L.pp_batch_set_stream.argtypes = [vp, vp]
    L.pp_batch_set_option.argtypes = [vp, i, ctypes.c_longlong]
    L.pp_batch_solve.argtypes = [vp]
    L.pp_batch_get_kernel_times.argtypes = [vp, fp, i, ip]
    L.pp_batch_solution.argtypes = [vp, i, ctypes.POINTER(vp)]
    L.pp_batch_get_rhs.argtypes = [vp, i, dp]
    ll = ctypes.c_longlong
    L.pp_server_create.argtypes = [ll, ll, i, ctypes.POINTER(vp)]
    L.pp_server_destroy.argtypes = [vp]
    L.pp_server_set_option.argtypes = [vp, i, ll]
    L.pp_server_submit.argtypes = [vp, i, i, i, ctypes.c_char_p, dp, i, dp, ctypes.c_ulonglong, i, ctypes.POINTER(ll)]
    L.pp_server_wait.argtypes = [vp, ll, dp, dp]
    L.pp_server_wait_ex.argtypes = [vp, ll, dp, dp, ctypes.POINTER(ll)]
    L.pp_server_footprint.argtypes = [vp, i, i, i, ctypes.c_char_p, dp, i, ctypes.POINTER(ll)]
    L.pp_server_stats.argtypes = [vp, ctypes.POINTER(ll), ctypes.POINTER(ll), ctypes.POINTER(ll), dp, ip]
    L.pp_batch_set_option_double.argtypes = [vp, i, ctypes.c_double]
    L.pp_copy_device.argtypes = [vp, vp, ctypes.c_longlong, i, vp]
    L.pp_batch_synchronize.argtypes = [vp]
    L.pp_init_trigonometric.argtypes = [vp, i]
    L.pp_get_error.argtypes = [vp, dp]
    L.pp_fill_trigonometric.argtypes = [vp, dp]
    L.pp_write_solution.argtypes = [ctypes.c_char_p, i, i, i, i, dp, dp, i]
    L.pp_hdf5_available.argtypes = []
    L.pp_solve_device_oop_noalloc.argtypes = [i, i, i, ctypes.c_char_p, dp, i, vp, vp, vp]
    for name in EXPORTED:
        if name != "pp_get_error_enum":
            getattr(L, name).restype = ctypes.c_int
    if path is None:
        _lib = L
    return L


def error_name(code: int) -> str:
    return load_library().pp_get_error_enum(int(code)).decode()


def _check(rc: int, where: str):
    if rc != SUCCESS:
        raise PittPackError(rc, where)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def size_supported(n: int) -> bool:
    return bool(load_library().pp_size_supported(int(n)))


def copy_device(d_in: int, d_out: int, n: int, mode: int = 0, stream: int = 0) -> None:
    """streaming device copy of n doubles (the achievable-bandwidth yardstick)"""
    _check(load_library().pp_copy_device(d_in, d_out, int(n), int(mode), stream or None), "copy_device")


def write_solution(folder: str, app: int, u, X, format: int = IO_HDF5) -> None:
    """Host-side solution writer behind PoissonHIP.IO (no GPU needed); u has shape (nz, ny, nx)."""
    f = np.ascontiguousarray(u, dtype=np.float64)
    nz, ny, nx = f.shape
    x = np.ascontiguousarray(X, dtype=np.float64)
    _check(load_library().pp_write_solution(os.fsencode(folder), int(app), nx, ny, nz, _dp(x), _dp(f), int(format)),
           "write_solution")


def hdf5_available() -> bool:
    return bool(load_library().pp_hdf5_available())


def selftest() -> int:
    return load_library().pp_selftest()


class PoissonHIP:
    """Python mirror of PencilDcmp's public solve interface on one MI355X.

    Same call order as src/main.cpp:73-118::

        M = PoissonHIP(nx, ny, nz)
        M.assignBoundary("NNNNDD"); M.setBox([0, 1, 0, 1, 0, 1]); M.setCoords(2)
        M.assignRhs(rhs); M.pittPack(); u = M.getSolution()

    ``rhs`` and the solution are real arrays of shape (nz, ny, nx) (x fastest), like the
    reference's ``assignRhs(double*)`` input.
    """

    def __init__(self, nx: int, ny: int, nz: int):
        self._L = load_library()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        h = ctypes.c_void_p()
        _check(self._L.pp_create(self.nx, self.ny, self.nz, ctypes.byref(h)), "pp_create")
        self._h = h

    # -- reference-named set-up -------------------------------------------------------
    def assignBoundary(self, bc) -> None:
        s = "".join(bc) if not isinstance(bc, str) else bc
        if len(s) != 6:
            raise PittPackError(INPUT_ARGS, "assignBoundary")
        _check(self._L.pp_assign_boundary(self._h, s.encode()), "assignBoundary")

    def setBox(self, X) -> None:
        x = np.ascontiguousarray(X, dtype=np.float64)
        _check(self._L.pp_set_box(self._h, _dp(x)), "setBox")

    def setCoords(self, dir: int = 2) -> None:
        _check(self._L.pp_set_coords(self._h, int(dir)), "setCoords")

    def assignRhs(self, rhs) -> None:
        f = np.ascontiguousarray(rhs, dtype=np.float64)
        if f.size != self.nx * self.ny * self.nz:
            raise PittPackError(INPUT_ARGS, "assignRhs")
        _check(self._L.pp_assign_rhs(self._h, _dp(f)), "assignRhs")

    def pittPack(self) -> None:
        _check(self._L.pp_pittpack(self._h), "pittPack")

    def getSolution(self) -> np.ndarray:
        u = np.empty((self.nz, self.ny, self.nx), dtype=np.float64)
        _check(self._L.pp_get_solution(self._h, _dp(u)), "getSolution")
        return u

    def IO(self, app: int = 0, folder: str = "soln", format: int | None = None) -> None:
        """PencilDcmp::IO (src/pencilDcmp.cpp:1827-1832): write <folder>/Pxdmf3d<app>.h5 (datasets
        X, Y, Z, Q as src/phdf5.cpp:1299-1808 lays them out) and its XDMF index
        <folder>/Pxdmf3d<app>.xmf; format=IO_RAW writes <folder>/Pxdmf3d<app>.bin instead."""
        if format is not None:
            self.set_option(OPT_IO_FORMAT, format)
        _check(self._L.pp_io(self._h, os.fsencode(folder), int(app)), "IO")

    def initializeTrigonometric(self, part: int = 0) -> None:
        """PencilDcmp::initializeTrigonometric (src/pencilDcmp.cpp:1871-2035): the reference's
        manufactured right-hand side (part 0 real, 1 imaginary part of its complex field)."""
        _check(self._L.pp_init_trigonometric(self._h, int(part)), "initializeTrigonometric")

    def getError(self) -> float:
        """PencilDcmp::getError (src/pencilDcmp.cpp:2040-2208): RMS error of the field against
        the manufactured case's exact solution (non-destructive)."""
        e = ctypes.c_double(0.0)
        _check(self._L.pp_get_error(self._h, ctypes.byref(e)), "getError")
        return e.value

    def fillTrigonometric(self) -> np.ndarray:
        """PencilDcmp::fillTrigonometric (src/pencilDcmp.cpp:5685-5765): -4 pi^2 cos(2 pi z)."""
        f = np.empty((self.nz, self.ny, self.nx), dtype=np.float64)
        _check(self._L.pp_fill_trigonometric(self._h, _dp(f)), "fillTrigonometric")
        return f

    # -- device / flow-solver forms ------------------------------------------------------
    def set_option(self, key: int, value: int) -> None:
        _check(self._L.pp_set_option(self._h, int(key), int(value)), "set_option")

    def set_option_double(self, key: int, value: float) -> None:
        _check(self._L.pp_set_option_double(self._h, int(key), float(value)), "set_option_double")

    def set_stream(self, stream_ptr: int) -> None:
        _check(self._L.pp_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "set_stream")

    def solve_device(self, d_ptr: int) -> None:
        """Solve in place on a device array of nx*ny*nz doubles (e.g. tensor.data_ptr())."""
        _check(self._L.pp_solve_device(self._h, ctypes.c_void_p(d_ptr)), "solve_device")

    def solve_device_oop(self, d_rhs: int, d_u: int) -> None:
        """RHS device array kept; solution written to d_u (may equal d_rhs)."""
        _check(self._L.pp_solve_device_oop(self._h, ctypes.c_void_p(d_rhs), ctypes.c_void_p(d_u)),
               "solve_device_oop")

    def stage_times(self):
        ms = (ctypes.c_float * 5)()
        _check(self._L.pp_get_stage_times(self._h, ms), "stage_times")
        return list(ms)

    def kernel_times(self):
        """Mean ms between consecutive launch marks of the recorded solves, in launch order."""
        ms = (ctypes.c_float * 64)()
        n = ctypes.c_int(0)
        _check(self._L.pp_get_kernel_times(self._h, ms, 64, ctypes.byref(n)), "kernel_times")
        return list(ms)[: n.value]

    def tuning(self) -> dict:
        """Configuration in use (after PP_OPT_AUTOTUNE: the tuner's choice)."""
        out = (ctypes.c_int * 9)()
        _check(self._L.pp_get_tuning(self._h, out, 9), "get_tuning")
        return {"tuned": bool(out[0]), "zsolver": out[1], "ztiles": out[2],
                "line_variant": {"x_forward": out[3], "y_forward": out[4], "y_inverse": out[5], "x_inverse": out[6]},
                "mixed_ztiles": out[7], "mixed_y_variant": out[8]}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value is not None:
            self._L.pp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Batch:
    """Batched form: many independent FP64 solves, one launch per stage and shape class.

        B = Batch()
        ids = [B.add(nx, ny, nz, bc, box) for ...]
        B.finalize()
        B.set_rhs(i, f_i) ...;  B.solve();  u_i = B.get_solution(i)
    """

    def __init__(self, precision: int = 0, refine: int | None = None, tau: float | None = None,
                 zero_mode: int | None = None, graph: int | None = None, autotune: int | None = None,
                 line_variant: int | None = None, zthomas: float | None = None, fuse: int | None = None):
        self._L = load_library()
        self._h = ctypes.c_void_p()
        _check(self._L.pp_batch_create(ctypes.byref(self._h)), "pp_batch_create")
        self.shapes = []
        if graph is not None:
            self.set_option(OPT_GRAPH, graph)
        if autotune is not None:
            self.set_option(OPT_AUTOTUNE, autotune)
        if line_variant is not None:
            self.set_option(OPT_LINE_VARIANT, line_variant)
        if precision:
            self.set_option(OPT_PRECISION, precision)
        if refine is not None:
            self.set_option(OPT_REFINE, refine)
        if tau is not None:
            _check(self._L.pp_batch_set_option_double(self._h, OPT_TAU, float(tau)), "batch_set_option_double")
        if zthomas is not None:
            _check(self._L.pp_batch_set_option_double(self._h, OPT_ZTHOMAS_TAU, float(zthomas)),
                   "batch_set_option_double")
        if zero_mode is not None:
            self.set_option(OPT_ZERO_MODE, zero_mode)
        if fuse is not None:
            self.set_option(OPT_FUSE_RESIDUAL, fuse)

    def add(self, nx: int, ny: int, nz: int, bc: str, box=None) -> int:
        X = np.ascontiguousarray(box if box is not None else [0, 1, 0, 1, 0, 1], dtype=np.float64)
        bcb = bc.encode() if isinstance(bc, str) else bytes(bc)
        i = ctypes.c_int(-1)
        _check(self._L.pp_batch_add(self._h, int(nx), int(ny), int(nz), bcb, _dp(X), ctypes.byref(i)), "batch_add")
        self.shapes.append((nz, ny, nx))
        return i.value

    def finalize(self) -> None:
        _check(self._L.pp_batch_finalize(self._h), "batch_finalize")

    def count(self):
        m, p = ctypes.c_int(0), ctypes.c_longlong(0)
        _check(self._L.pp_batch_count(self._h, ctypes.byref(m), ctypes.byref(p)), "batch_count")
        return m.value, p.value

    def launch_count(self):
        a = (ctypes.c_int * 3)()
        _check(self._L.pp_batch_launch_count(self._h, a), "batch_launch_count")
        return list(a)

    def field_ptr(self, i: int) -> int:
        """device pointer of member i's right-hand side"""
        p = ctypes.c_void_p()
        _check(self._L.pp_batch_field(self._h, int(i), ctypes.byref(p)), "batch_field")
        return p.value

    def solution_ptr(self, i: int) -> int:
        p = ctypes.c_void_p()
        _check(self._L.pp_batch_solution(self._h, int(i), ctypes.byref(p)), "batch_solution")
        return p.value

    def set_rhs(self, i: int, rhs) -> None:
        f = np.ascontiguousarray(rhs, dtype=np.float64)
        if f.shape != self.shapes[i]:
            raise ValueError(f"rhs shape {f.shape} != {self.shapes[i]}")
        _check(self._L.pp_batch_set_rhs(self._h, int(i), _dp(f)), "batch_set_rhs")

    def get_solution(self, i: int) -> np.ndarray:
        u = np.empty(self.shapes[i], dtype=np.float64)
        _check(self._L.pp_batch_get_solution(self._h, int(i), _dp(u)), "batch_get_solution")
        return u

    def get_rhs(self, i: int) -> np.ndarray:
        f = np.empty(self.shapes[i], dtype=np.float64)
        _check(self._L.pp_batch_get_rhs(self._h, int(i), _dp(f)), "batch_get_rhs")
        return f

    def fill_uniform(self, seed: int, round_f32: bool = False) -> None:
        _check(self._L.pp_batch_fill_uniform(self._h, int(seed), int(bool(round_f32))), "batch_fill_uniform")

    def set_stream(self, stream_ptr: int) -> None:
        _check(self._L.pp_batch_set_stream(self._h, ctypes.c_void_p(stream_ptr)), "batch_set_stream")

    def set_option(self, key: int, value: int) -> None:
        _check(self._L.pp_batch_set_option(self._h, int(key), int(value)), "batch_set_option")

    def solve(self) -> None:
        _check(self._L.pp_batch_solve(self._h), "batch_solve")

    def tuning(self) -> int:
        """(line class, direction) pairs running the alternative launch shape (autotuned)."""
        n = ctypes.c_int(0)
        _check(self._L.pp_batch_tuning(self._h, ctypes.byref(n)), "batch_tuning")
        return n.value

    def synchronize(self) -> None:
        _check(self._L.pp_batch_synchronize(self._h), "batch_synchronize")

    def kernel_times(self):
        ms = (ctypes.c_float * 64)()
        n = ctypes.c_int(0)
        _check(self._L.pp_batch_get_kernel_times(self._h, ms, 64, ctypes.byref(n)), "batch_kernel_times")
        return list(ms)[: n.value]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value is not None:
            self._L.pp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Serving form: submit solves from any thread, wait for them; batches run concurrently.

        S = Server(hbm_bytes=8 << 30, staging_bytes=2 << 30, streams=4)
        t = S.submit((nz, ny, nx), "PPPPNN", rhs=f)        # or seed=... for device-generated RHS
        u, latency_ms = S.wait(t)
    """

    def __init__(self, hbm_bytes: int = 8 << 30, staging_bytes: int = 2 << 30, streams: int = 4):
        self._L = load_library()
        self._h = ctypes.c_void_p()
        _check(self._L.pp_server_create(int(hbm_bytes), int(staging_bytes), int(streams), ctypes.byref(self._h)),
               "pp_server_create")
        self._pending = {}

    def set_option(self, key: int, value: int) -> None:
        _check(self._L.pp_server_set_option(self._h, int(key), int(value)), "server_set_option")

    def submit(self, shape, bc: str, box=None, precision: int = 0, rhs=None, seed: int = 0,
               want_solution: bool = True) -> int:
        nz, ny, nx = shape
        X = np.ascontiguousarray(box if box is not None else [0, 1, 0, 1, 0, 1], dtype=np.float64)
        f = None
        if rhs is not None:
            f = np.ascontiguousarray(rhs, dtype=np.float64)
            if f.shape != (nz, ny, nx):
                raise ValueError("rhs shape")
        t = ctypes.c_longlong(0)
        rp = _dp(f) if f is not None else None
        _check(self._L.pp_server_submit(self._h, int(nx), int(ny), int(nz), bc.encode(), _dp(X), int(precision), rp,
                                        int(seed), int(bool(want_solution)), ctypes.byref(t)), "server_submit")
        # keep the host buffers alive until the request is collected
        self._pending[t.value] = (f, X, (nz, ny, nx), bool(want_solution))
        return t.value

    def wait(self, ticket: int, with_batch: bool = False, out: np.ndarray | None = None):
        """(solution or None, latency ms[, launch number of the batch that ran it]); `out`: a
        C-contiguous float64 buffer of at least the request's size to receive the solution
        (returned as a view of it), e.g. one reused across requests"""
        f, X, shape, want = self._pending.pop(ticket)
        u = None
        if want:
            n = shape[0] * shape[1] * shape[2]
            if out is not None:
                if out.dtype != np.float64 or not out.flags.c_contiguous or out.size < n:
                    raise ValueError("out: C-contiguous float64 buffer of the request's size needed")
                u = out.reshape(-1)[:n].reshape(shape)
            else:
                u = np.empty(shape, dtype=np.float64)
        lat = ctypes.c_double(0)
        seq = ctypes.c_longlong(-1)
        _check(self._L.pp_server_wait_ex(self._h, int(ticket), _dp(u) if u is not None else None, ctypes.byref(lat),
                                         ctypes.byref(seq)), "server_wait")
        return (u, lat.value, seq.value) if with_batch else (u, lat.value)

    def footprint(self, shape, bc: str, box=None, precision: int = 0) -> int:
        """device bytes a request of this shape takes in the pool"""
        nz, ny, nx = shape
        X = np.ascontiguousarray(box if box is not None else [0, 1, 0, 1, 0, 1], dtype=np.float64)
        b = ctypes.c_longlong(0)
        _check(self._L.pp_server_footprint(self._h, int(nx), int(ny), int(nz), bc.encode(), _dp(X), int(precision),
                                           ctypes.byref(b)), "server_footprint")
        return b.value

    def stats(self):
        c, p, b = ctypes.c_longlong(0), ctypes.c_longlong(0), ctypes.c_longlong(0)
        span, mc = ctypes.c_double(0), ctypes.c_int(0)
        _check(self._L.pp_server_stats(self._h, ctypes.byref(c), ctypes.byref(p), ctypes.byref(b), ctypes.byref(span),
                                       ctypes.byref(mc)), "server_stats")
        return {"completed": c.value, "points": p.value, "batches": b.value, "span_ms": span.value,
                "max_concurrent_batches": mc.value}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value is not None:
            self._L.pp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def solve(rhs, bc: str, box=None, zero_mode: int | None = None, zsolver: int | None = None,
          precision: int = 0, refine: int | None = None, tau: float | None = None,
          line_variant: int | None = None, autotune: int | None = None, zlayout: int | None = None,
          fuse: int | None = None) -> np.ndarray:
    """One-shot solve of an (nz, ny, nx) FP64 RHS (precision 0: FP64; 1: mixed + refinement).
    zlayout: 1 (default) hands the z solve the y transform's z tiles where it can, 0 in place.
    fuse: mixed form, 1 (default) computes each correction's residual inside its x forward, 0
    runs the residual kernel and the transform apart."""
    f = np.ascontiguousarray(rhs, dtype=np.float64)
    nz, ny, nx = f.shape
    with PoissonHIP(nx, ny, nz) as M:
        if zero_mode is not None:
            M.set_option(OPT_ZERO_MODE, zero_mode)
        if line_variant is not None:
            M.set_option(OPT_LINE_VARIANT, line_variant)
        if autotune is not None:
            M.set_option(OPT_AUTOTUNE, autotune)
        if zsolver is not None:
            M.set_option(OPT_ZSOLVER, zsolver)
        if zlayout is not None:
            M.set_option(OPT_ZLAYOUT, zlayout)
        if fuse is not None:
            M.set_option(OPT_FUSE_RESIDUAL, fuse)
        if precision:
            M.set_option(OPT_PRECISION, precision)
        if refine is not None:
            M.set_option(OPT_REFINE, refine)
        if tau is not None:
            M.set_option_double(OPT_TAU, tau)
        M.assignBoundary(bc)
        M.setBox(box if box is not None else [0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        M.assignRhs(f)
        M.pittPack()
        return M.getSolution()
