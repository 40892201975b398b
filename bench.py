#!/usr/bin/env python3
"""Headline benchmark of the MI355X PittPack solve path.

Workload (one GPU): the plain FP64 drop-in at 512^3 -- Dirichlet-Dirichlet in x,
Neumann-Dirichlet in y, Dirichlet-Dirichlet in z, box [0,1]^3, RHS U[-1,1] FP64 from the
counter generator with seed 0.  One step = one full solve (x forward, y forward, z solve,
y inverse, x inverse) of the resident RHS into a resident solution array.

Prints one JSON line: grid points solved per second, the roofline of the dominant kernel
(HIP events on the solver's stream inside the timed region) and a CPU baseline (the
oracle, i.e. the C restatement of the reference algorithm, on a bounded 256^3 sample of
the same workload, one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BC = "DDNDDD"
STAGES = ["x_forward", "y_forward", "z_solve", "y_inverse", "x_inverse"]


def cpu_baseline(n: int = 256, bc: str = BC):
    import numpy as np
    import oracle
    from pittpack_amd import rng

    f = rng.uniform_pm1((n, n, n), seed=0)
    t0 = time.perf_counter()
    oracle.solve(f, bc)
    dt = time.perf_counter() - t0
    return {
        "value": n ** 3 / dt,
        "unit": "grid points/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one FP64 solve at {n}^3, bc {bc}, seed-0 U[-1,1] RHS, oracle (C restatement "
                  f"of the reference algorithm), single thread: {dt:.2f} s",
    }


def load_traffic(name: str):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        try:
            with open(p) as fh:
                return json.load(fh)
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zsolver", type=int, default=None, help="0 PCR->Thomas, 1 partitioned Thomas")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import torch

    import pittpack_amd as pa
    from pittpack_amd import rng

    n = args.n
    dev = torch.device("cuda:0")
    f_host = rng.uniform_pm1((n, n, n), seed=0)
    f = torch.from_numpy(f_host).to(dev)
    u = torch.empty_like(f)
    del f_host

    M = pa.PoissonHIP(n, n, n)
    M.assignBoundary(BC)
    M.setBox([0, 1, 0, 1, 0, 1])
    M.setCoords(2)
    if args.zsolver is not None:
        M.set_option(pa.OPT_ZSOLVER, args.zsolver)
    stream = torch.cuda.current_stream(dev)
    M.set_stream(stream.cuda_stream)

    for _ in range(args.warmup):
        M.solve_device_oop(f.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    M.set_option(pa.OPT_TIMING, max(1, args.steps))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        M.solve_device_oop(f.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    stage_ms = M.stage_times()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    pts = float(n) ** 3
    value = pts / (ms_per_step * 1e-3)

    bytes_per_stage = 16.0 * pts  # FP64: each stage reads and writes the field once
    dom = int(np.argmax(stage_ms))
    achieved = bytes_per_stage / (stage_ms[dom] * 1e-3) / 1e9
    traffic = load_traffic("traffic_fp64_512.json")
    tr_val = None
    if traffic and STAGES[dom] in traffic:
        tr_val = traffic[STAGES[dom]]

    # device copy bandwidth of the same array, for reference
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    for _ in range(5):
        u.copy_(f)
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * pts * 8 / (time.perf_counter() - c0) / 1e9

    M.close()

    out = {
        "metric": "grid points solved/sec (FP64 drop-in, 512^3)",
        "value": value,
        "unit": "grid points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain FP64 drop-in, {n}^3, bc x DD / y ND / z DD, box [0,1]^3, "
                        "RHS U[-1,1] seed 0 (counter generator)",
            "n": n,
            "bc": BC,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": STAGES[dom],
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": tr_val,
        },
        "stages_ms": dict(zip(STAGES, stage_ms)),
        "stages_gbs": {k: bytes_per_stage / (v * 1e-3) / 1e9 for k, v in zip(STAGES, stage_ms)},
        "copy_gbs": copy_gbs,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
