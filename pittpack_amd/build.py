"""Build the HIP library pittpack_amd/lib/libpittpack_hip.so for gfx950 with hipcc.

Each translation unit under csrc/ is compiled to an object in parallel, then linked.
Usage: python -m pittpack_amd.build   (or pittpack_amd.build.build() from Python)
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "libpittpack_hip.so")
# the C++ binding of include/pittpack_hip.hpp, driven in main.cpp's call order (tests load it)
CXX_SRC = os.path.join(HERE, "cxx", "pp_cxx_check.cpp")
CXX_OUT = os.path.join(OUT_DIR, "libpp_cxx_check.so")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # packed FP32 instructions off: some of their forms are unreliable on these parts
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src: str, extra) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-I" + os.path.join(ROOT, "include"), "-I" + SRC,
           "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    with open(obj + ".log", "w") as fh:
        fh.write(res.stdout + res.stderr)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def _fresh(src: str, newest_header: float) -> bool:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    return os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header)


def build(verbose: bool = False, extra=(), jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every translation unit (incremental=True: only those older than their source
    or than any header) and link the library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(SRC, "*.hip")))
    headers = glob.glob(os.path.join(SRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    newest = max((os.path.getmtime(h) for h in headers), default=0.0)
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))

    def one(s):
        if incremental and not extra and _fresh(s, newest):
            return os.path.join(OBJ_DIR, os.path.basename(s) + ".o")
        return _compile(s, list(extra))

    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(one, srcs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    build_cxx_check()
    return OUT


def build_cxx_check() -> str:
    """g++ build of the C++ binding against the library (rpath: its own folder)."""
    tmp = CXX_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"), CXX_SRC,
           "-L" + OUT_DIR, "-lpittpack_hip", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("g++ failed on the C++ binding")
    os.replace(tmp, CXX_OUT)
    return CXX_OUT


if __name__ == "__main__":
    args = sys.argv[1:]
    inc = "--incremental" in args
    print(build(verbose=True, extra=[a for a in args if a != "--incremental"], incremental=inc))
