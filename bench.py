#!/usr/bin/env python3
"""Headline benchmark of the MI355X PittPack solve path.

Workload (one GPU): the plain FP64 drop-in at 512^3 -- Dirichlet-Dirichlet in x,
Neumann-Dirichlet in y, Dirichlet-Dirichlet in z, box [0,1]^3, RHS U[-1,1] FP64 from the
counter generator with seed 0.  One step = one full solve (x forward, y forward, z solve,
y inverse, x inverse) of the resident RHS into a resident solution array.

The same line carries the mixed form on the same workload (FP32 transforms and z solves,
FP64 near-singular columns, `--refine` FP64-residual refinement steps): its throughput, its
ratio to the FP64 form and its error against the FP64 result.

Prints one JSON line: grid points solved per second, the roofline of the dominant kernel
(HIP events recorded on the solver's stream between its launches, inside the timed region),
the achievable bandwidth of a streaming copy of the same array, the 1024^3 FP64 and mixed
runs (the mixed one against the FP64 one), the batched form (a stratified sample of members
against the oracle), the serving form at several arrival rates and closed loop, and a CPU
baseline (the oracle, i.e. the C restatement of the reference algorithm, on 10 solves of the
same 512^3 workload on the host's cores through OpenMP, compared with the GPU solution).
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

# launch order of one solve and the algorithmic bytes per grid point of each kernel
FP64_KERNELS = [("x_forward", 16), ("y_forward", 16), ("z_solve", 16), ("y_inverse", 16), ("x_inverse", 16)]
def progress(msg: str):
    """one line per finished phase on stderr (the JSON result goes to stdout)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


MIXED_INIT = [("x_forward_f64_to_f32", 12), ("y_forward_f32", 8), ("z_solve_f32", 8), ("z_fix_f64", 0),
              ("y_inverse_f32", 8)]
MIXED_STEP = [("x_forward_f32", 8), ("y_forward_f32", 8), ("z_solve_f32", 8), ("z_fix_f64", 0), ("y_inverse_f32", 8)]


def kernel_table(precision: int, refine: int, fused: bool = True):
    """launch order of one solve with the algorithmic bytes per grid point of each kernel.  The
    mixed form holds the solution as FP32 parts while it is refined (u0, then the sum of the
    corrections before the last); only the last store writes u in FP64.  fused: each
    correction's residual is computed inside its x-forward transform (one kernel reading f, u0
    (+ the correction sum) and writing the transformed lines), else the residual kernel writes
    r and the transform reads it."""
    if precision == 0:
        return FP64_KERNELS
    out = list(MIXED_INIT)
    out.append(("x_inverse_f32", 8) if refine > 0 else ("x_inverse_final_f64", 12))  # c -> u0 (or u)
    for k in range(1, refine + 1):
        if fused:
            out.append((f"residual_x_forward_u0.{k}", 16) if k == 1 else (f"residual_x_forward_u0_csum.{k}", 20))
            out += [(f"{name}.{k}", b) for name, b in MIXED_STEP[1:]]
        else:
            # f, u0 (+ the correction sum), r
            out.append((f"residual_u0.{k}", 16) if k == 1 else (f"residual_u0_csum.{k}", 20))
            out += [(f"{name}.{k}", b) for name, b in MIXED_STEP]
        if k < refine:
            out.append((f"x_inverse_f32.{k}", 8 if k == 1 else 12))  # c (+ sum) -> sum
        else:
            out.append((f"x_inverse_final_f64.{k}", 16 if k == 1 else 20))  # c, u0 (+ sum) -> u
    return out


def host_cores():
    """(physical cores of the host from lscpu, logical CPUs this process may run on)"""
    import subprocess

    phys = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        phys = int(kv["Core(s) per socket"].strip()) * int(kv["Socket(s)"].strip())
    except Exception:
        pass
    return phys, len(os.sched_getaffinity(0))


def cpu_baseline(u_gpu=None, n: int = 512, bc: str = BC, solves: int = 10):
    """The oracle (C restatement of the reference algorithm) on the same 512^3 workload, timed
    over `solves` solves into a preallocated workspace (the reference's window averages 10
    solves, src/pencilDcmp.cpp:5900,6332-6376), with the FFT plan made and destroyed per line
    as the reference's CPU path does (src/poissonCPU.cpp:52-63); three plan-once solves are
    timed beside it.  u_gpu: the GPU FP64 solution of the same RHS, compared with the
    oracle's."""
    import numpy as np

    import oracle
    from pittpack_amd import rng

    f = rng.uniform_pm1((n, n, n), seed=0)
    w = oracle.workspace(f.shape)
    u = np.zeros_like(f)
    oracle.set_per_line_plans(True)
    t0 = time.perf_counter()
    for _ in range(solves):
        oracle.solve(f, bc, zero_mode=1, work=w, out=u)
    dt = (time.perf_counter() - t0) / solves
    oracle.set_per_line_plans(False)
    t0 = time.perf_counter()
    for _ in range(3):
        oracle.solve(f, bc, zero_mode=1, work=w, out=u)
    dt_once = (time.perf_counter() - t0) / 3
    threads = oracle.threads()
    phys, logical = host_cores()
    out = {
        "value": n ** 3 / dt,
        "unit": "grid points/s",
        "cores": threads,
        "host_physical_cores": phys,
        "host_logical_cpus": logical,
        "kind": "port",
        "sample": f"{solves} FP64 solves at {n}^3, bc {bc}, seed-0 U[-1,1] RHS, oracle (C restatement of "
                  f"the reference algorithm, FFT plan made and destroyed per line as poissonCPU.cpp does; "
                  f"lines, planes and z columns split over {threads} OpenMP threads, as the reference splits "
                  f"them over ranks; workspace preallocated): {dt:.2f} s per solve",
        "plan_once_value": n ** 3 / dt_once,
    }
    if u_gpu is not None:
        d = np.abs(u_gpu - u)
        out["gpu_fp64_rel_max_diff_vs_oracle"] = float(d.max() / np.abs(u).max())
        out["gpu_fp64_rel_l2_diff_vs_oracle"] = float(np.sqrt((d ** 2).sum() / (u ** 2).sum()))
    return out


XY_COMBOS = ["PPPP", "NNNN", "DDDD", "NNDD"]  # x/y faces of the batched and serving workloads
Z_COMBOS = ["DD", "NN"]
XY_SIDES = [32, 48, 64, 96, 128, 192]


def stratified_members(members, per_class: int = 8):
    """indices of up to per_class members from each boundary combination, spread over the
    class's sizes (smallest to largest)"""
    by_bc = {}
    for i, m in enumerate(members):
        by_bc.setdefault(m[3], []).append(i)
    picks = []
    for bc in sorted(by_bc):
        idx = sorted(by_bc[bc], key=lambda i: members[i][0] * members[i][1] * members[i][2])
        step = max(1, len(idx) // per_class)
        picks += idx[::step][:per_class]
    return sorted(picks)


def bench_batched(pa, count: int, steps: int, refine: int, sample: int = 8, autotune: bool = True):
    """4096 independent solves (nx, ny from {32..192}, nz in [16, 256], BC uniform over four
    x/y combos x z in {DD, NN}, RHS U[-1,1] rounded to FP32 from seed 2), mixed form with
    `refine` steps; the same batch in FP64 gives the error reference and its own rate.  A
    stratified sample (8 members of each of the 8 boundary combinations) of both forms is
    checked against the oracle."""
    import numpy as np

    import oracle

    r = np.random.default_rng(2)
    members = []
    for _ in range(count):
        nx, ny = int(r.choice(XY_SIDES)), int(r.choice(XY_SIDES))
        nz = int(r.integers(16, 257))
        bc = XY_COMBOS[int(r.integers(4))] + Z_COMBOS[int(r.integers(2))]
        members.append((nx, ny, nz, bc))
    picks = stratified_members(members)
    out = {}
    ref = {}
    oracle_sol = {}
    for prec in (0, 1):
        # FP64: line classes tuned at finalize (before the right-hand sides are written)
        B = pa.Batch(precision=prec, refine=refine if prec else None, autotune=int(autotune and prec == 0))
        for nx, ny, nz, bc in members:
            B.add(nx, ny, nz, bc)
        t0 = time.perf_counter()
        B.finalize()
        plan_s = time.perf_counter() - t0
        B.fill_uniform(2, round_f32=True)
        # first solve eager, second captured into a HIP graph; the timed solves replay it
        for _ in range(2):
            B.solve()
        B.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            B.solve()
        B.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        # per-stage times from events between the launches (eager solves)
        B.set_option(pa.OPT_TIMING, steps)
        for _ in range(steps):
            B.solve()
        kt = B.kernel_times()
        m, pts = B.count()
        launches = B.launch_count()
        key = "fp64" if prec == 0 else "mixed"
        sols = {i: B.get_solution(i) for i in picks}
        errs_o = []
        for i in picks:
            nx, ny, nz, bc = members[i]
            if prec == 0:
                oracle_sol[i] = oracle.solve(B.get_rhs(i), bc, zero_mode=1)
            o = oracle_sol[i]
            errs_o.append(float(np.abs(sols[i] - o).max() / np.abs(o).max()))
        if prec == 0:
            ref = sols
        res = {"ms_per_step": ms, "value": pts / (ms * 1e-3), "launches_per_stage_xyz": launches,
               "plan_and_upload_s": plan_s, "hip_graph": True, "kernels_ms": kt}
        res["oracle_checked_members"] = len(picks)
        res["rel_max_diff_vs_oracle_max"] = max(errs_o)
        if prec == 0:
            res["alternative_shapes"] = B.tuning()
        if prec == 1:
            errs = [float(np.abs(sols[i] - ref[i]).max() / np.abs(ref[i]).max()) for i in picks]
            res["rel_max_diff_vs_fp64_sampled"] = max(errs)
            res["refine"] = refine
        out[key] = res
        out["members"], out["points"] = m, pts
        B.close()
    out["workload"] = (f"{count} solves, nx,ny in {XY_SIDES}, nz in [16,256], bc x/y in {XY_COMBOS} x z in "
                       f"{Z_COMBOS}, RHS U[-1,1] rounded to FP32, seed 2 (+member index)")
    return out


def serving_requests(n_req: int, seed: int = 5):
    """request list: sides log-uniform integers in [32, 512] per axis, BC uniform over four x/y
    combos x z in {DD, NN}, FP64 or mixed 50/50"""
    import numpy as np

    r = np.random.default_rng(seed)
    reqs = []
    for i in range(n_req):
        side = np.exp(r.uniform(np.log(32), np.log(512), 3))
        nx, ny, nz = (int(round(v)) for v in side)
        bc = XY_COMBOS[int(r.integers(4))] + Z_COMBOS[int(r.integers(2))]
        reqs.append(((nz, ny, nx), bc, int(r.integers(2)), 7000 + i))
    return reqs


def bench_serving(pa, n_req: int, rate: float, budget_gb: float, streams: int, reqs=None, collectors: int = 4,
                  to_host: bool = True):
    """Poisson arrivals at `rate`/s (rate <= 0: closed loop, every request submitted at once);
    RHS from the device counter generator.  to_host: every solution is copied back to the host
    (pinned staging, then into a buffer each of `collectors` waiting threads reuses); else the
    solutions stay on the device and a request completes when its batch has finished."""
    import threading

    import numpy as np

    reqs = reqs or serving_requests(n_req)
    n_req = len(reqs)
    r = np.random.default_rng(6)
    gaps = r.exponential(1.0 / rate, n_req) if rate > 0 else np.zeros(n_req)
    S = pa.Server(hbm_bytes=int(budget_gb * (1 << 30)), staging_bytes=16 << 30, streams=streams)
    tickets = [None] * n_req
    ready = [threading.Event() for _ in range(n_req)]
    lat = [0.0] * n_req
    biggest = max(s[0] * s[1] * s[2] for s, _, _, _ in reqs)

    def collector(k):
        buf = np.ones(biggest) if to_host else None  # touched once: no page faults per request
        for i in range(k, n_req, collectors):
            ready[i].wait()
            _, lat[i] = S.wait(tickets[i], out=buf)

    ths = [threading.Thread(target=collector, args=(k,)) for k in range(collectors)]
    for th in ths:
        th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i, (shape, bc, prec, seed) in enumerate(reqs):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = S.submit(shape, bc, precision=prec, seed=seed, want_solution=to_host)
        ready[i].set()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    st = S.stats()
    S.close()
    pts = sum(s[0] * s[1] * s[2] for s, _, _, _ in reqs)
    lat_a = np.array(lat)
    return {
        "requests": n_req, "rate_per_s": rate if rate > 0 else "closed loop", "hbm_budget_gb": budget_gb,
        "streams": streams,
        "p50_ms": float(np.percentile(lat_a, 50)), "p99_ms": float(np.percentile(lat_a, 99)),
        "mean_ms": float(lat_a.mean()), "solves_per_s": n_req / wall, "points_per_s": pts / wall,
        "batches": st["batches"], "requests_per_batch": n_req / max(1, st["batches"]),
        "max_concurrent_batches": st["max_concurrent_batches"],
        "mean_points_per_request": pts / n_req,
        "solutions": "copied to host" if to_host else "left on the device",
    }


def bench_serving_sweep(pa, n_req: int, rates, budget_gb: float = 200.0, streams: int = 4):
    """the serving workload at several arrival rates and closed loop (same request list)"""
    reqs = serving_requests(n_req)
    runs = [bench_serving(pa, n_req, rate, budget_gb, streams, reqs) for rate in rates]
    # the engine alone: the same requests at the highest rates with the solutions left on the
    # device (the host copy of ~35 MB per request bounds the runs above)
    runs += [bench_serving(pa, n_req, rate, budget_gb, streams, reqs, to_host=False)
             for rate in rates if rate <= 0 or rate >= 1000]
    return {
        "workload": ("Poisson arrivals; sides log-uniform integers in [32,512] per axis (any length), "
                     f"bc x/y in {XY_COMBOS} x z in {Z_COMBOS}, FP64 / mixed 50/50, device-generated RHS; "
                     "latency = submit to solution in pinned host memory (or, solutions left on the "
                     "device, to the end of the request's batch)"),
        "runs": runs,
    }


def bench_large(pa, torch, n: int, precision: int, refine: int, steps: int, ref=None):
    """FP64 form (autotuned) or mixed form at n^3, periodic x/y, Neumann z, FP32 Gaussian RHS
    (torch generator, seed 1) with the volume mean removed; FP64 discrete residual of the
    result and, given the FP64 form's solution `ref`, the difference to it.  Returns
    (record, solution)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(1)
    f = torch.randn((n, n, n), generator=g, device=dev, dtype=torch.float32).double()
    f -= f.mean()
    u = torch.empty_like(f)
    M = pa.PoissonHIP(n, n, n)
    M.assignBoundary("PPPPNN")
    M.setBox([0, 1, 0, 1, 0, 1])
    M.setCoords(2)
    if precision == 1:
        M.set_option(pa.OPT_PRECISION, 1)
        M.set_option(pa.OPT_REFINE, refine)
    M.set_option(pa.OPT_AUTOTUNE, 1)  # tuned on the first (untimed) solve
    M.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    M.solve_device_oop(f.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        M.solve_device_oop(f.data_ptr(), u.data_ptr())
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    M.close()
    # FP64 residual f h^2 - L u (P wrap in x and y, N mirror in z), relative to max(|f h^2|, |u|)
    r = f * (1.0 / n) ** 2
    scale = max(float(r.abs().max()), float(u.abs().max()))
    for ax in (1, 2):
        r -= torch.roll(u, -1, ax) + torch.roll(u, 1, ax) - 2.0 * u
    up = torch.cat([u[1:], u[-1:]], 0)
    um = torch.cat([u[:1], u[:-1]], 0)
    r -= up + um - 2.0 * u
    res = float(r.abs().max()) / scale
    del f, r, up, um
    form = f"mixed form, {refine} refinement steps, autotuned" if precision == 1 else "FP64 form, autotuned"
    rec = {"workload": f"{form}, {n}^3, bc PPPPNN, FP32 N(0,1) RHS (torch generator seed 1), mean removed",
           "ms_per_step": ms, "value": n ** 3 / (ms * 1e-3), "fp64_residual_rel": res}
    if ref is not None:
        d = u - ref
        rec["rel_max_diff_vs_fp64"] = float(d.abs().max() / ref.abs().max())
        rec["rel_l2_diff_vs_fp64"] = float(d.norm() / ref.norm())
        del d
    torch.cuda.empty_cache()
    return rec, u


def load_traffic(workload_key: str, kernel: str):
    """HBM bytes per launch of `kernel`, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes (scaled by the copy-kernel calibration), or None."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as fh:
            t = json.load(fh)
        return t[workload_key][kernel]["bytes_per_launch"]
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--refine", type=int, default=2, help="refinement steps of the mixed form")
    ap.add_argument("--no-mixed", action="store_true", help="time the FP64 form only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--zsolver", type=int, default=None,
                    help="0 PCR->Thomas, 1 partitioned Thomas, 2 multigrid, 3 full PCR, 4 cyclic reduction")
    ap.add_argument("--no-autotune", action="store_true", help="FP64 and mixed forms in the fixed default configuration")
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel of every solve (no HIP graph)")
    ap.add_argument("--batch-count", type=int, default=4096, help="solves in the batched workload (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=300, help="requests per serving run (0: skip)")
    ap.add_argument("--serve-rates", type=str, default="100,1000,3000,0",
                    help="Poisson arrival rates per second of the serving sweep (0: closed loop)")
    ap.add_argument("--large", type=int, default=1024, help="side of the large FP64 and mixed runs (0: skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only (no multi-GPU path); use --gpus 1"}))
        sys.exit(2)

    import torch

    import pittpack_amd as pa
    from pittpack_amd import rng

    n = args.n
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)  # default priority; graphs need a non-null stream
    torch.cuda.set_stream(stream)
    f_host = rng.uniform_pm1((n, n, n), seed=0)
    f = torch.from_numpy(f_host).to(dev)
    u = torch.empty_like(f)
    del f_host

    pts = float(n) ** 3

    def timed_form(precision: int):
        """warmup + `steps` timed solves of one form; returns (ms/solve, per-kernel ms, handle)"""
        M = pa.PoissonHIP(n, n, n)
        M.assignBoundary(BC)
        M.setBox([0, 1, 0, 1, 0, 1])
        M.setCoords(2)
        if args.zsolver is not None:
            M.set_option(pa.OPT_ZSOLVER, args.zsolver)
        if not args.no_autotune:
            M.set_option(pa.OPT_AUTOTUNE, 1)  # tuned on the first (warm-up) solve
        if precision:
            M.set_option(pa.OPT_PRECISION, precision)
            M.set_option(pa.OPT_REFINE, args.refine)
        M.set_stream(stream.cuda_stream)
        # repeated solves on the same pointers replay a captured HIP graph (the first warm-up
        # solve runs eagerly -- and tunes the FP64 form -- the second is captured)
        M.set_option(pa.OPT_GRAPH, 0 if args.no_graph else 1)
        for _ in range(max(2, args.warmup)):
            M.solve_device_oop(f.data_ptr(), u.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            M.solve_device_oop(f.data_ptr(), u.data_ptr())
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # per-kernel durations: HIP events between the launches of `steps` more (eager) solves
        M.set_option(pa.OPT_TIMING, max(1, args.steps))
        for _ in range(args.steps):
            M.solve_device_oop(f.data_ptr(), u.data_ptr())
        torch.cuda.synchronize()
        return (t1 - t0) * 1e3 / args.steps, M.kernel_times(), M

    def roofline(precision: int, kt):
        # the mixed form fuses each correction's residual into its x forward where the grid
        # allows (one launch mark fewer per step)
        table = kernel_table(precision, args.refine, fused=True)
        if len(kt) != len(table):
            table = kernel_table(precision, args.refine, fused=False)
        names = [k for k, _ in table]
        kms = dict(zip(names, kt))
        bytes_pp = dict(table)
        dom = max((k for k in names if bytes_pp[k] > 0), key=lambda k: kms[k])
        alg_bytes = bytes_pp[dom] * pts
        achieved = alg_bytes / (kms[dom] * 1e-3) / 1e9
        wl_key = f"{'fp64' if precision == 0 else 'mixed'}_{BC}_{n}"
        roof = {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(wl_key, dom.split(".")[0]),
            "alg_bytes_per_launch": alg_bytes,
        }
        gbs = {k: bytes_pp[k] * pts / (kms[k] * 1e-3) / 1e9 for k in names if bytes_pp[k] > 0}
        return roof, kms, gbs

    # the headline: the plain FP64 drop-in
    ms_per_step, kt, M = timed_form(0)
    value = pts / (ms_per_step * 1e-3)
    roof, kms, gbs = roofline(0, kt)
    tuning = M.tuning()
    u64 = u.clone()
    M.close()
    progress(f"FP64 {n}^3: {ms_per_step:.3f} ms per solve")

    # the mixed form on the same workload, its throughput ratio and its error vs the FP64 form
    mixed = None
    if not args.no_mixed:
        mms, mkt, M = timed_form(1)
        mroof, mkms, mgbs = roofline(1, mkt)
        diff = (u - u64)
        mixed = {
            "value": pts / (mms * 1e-3),
            "ms_per_step": mms,
            "refine": args.refine,
            "vs_fp64": ms_per_step / mms,
            "rel_max_diff_vs_fp64": float(diff.abs().max() / u64.abs().max()),
            "rel_l2_diff_vs_fp64": float(diff.norm() / u64.norm()),
            "roofline": mroof,
            "kernels_ms": mkms,
            "tuning": {k: v for k, v in M.tuning().items() if k.startswith("mixed")},
        }
        del diff
        M.close()
        progress(f"mixed {n}^3: {mms:.3f} ms per solve")
    # host copy of the FP64 solution for the comparison with the oracle (CPU baseline leg)
    u64_host = u64.cpu().numpy() if not args.no_cpu_baseline else None
    del u64

    # achievable bandwidth on the same array (8 B read + 8 B written per point): torch's
    # elementwise kernel and the library's streaming copy (cached / nontemporal), HIP events
    # on the stream the copies run on
    def copy_rate(fn, reps=20):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return reps * 2 * pts * 8 / (e0.elapsed_time(e1) * 1e-3) / 1e9

    copy_gbs = copy_rate(lambda: torch.neg(f, out=u))
    stream_copy = {m: copy_rate(lambda: pa.copy_device(f.data_ptr(), u.data_ptr(), int(pts), mode,
                                                        stream.cuda_stream))
                   for m, mode in (("cached", 0), ("nontemporal", 1), ("nontemporal_stores", 2),
                                   ("nontemporal_loads", 3))}
    M = None

    out = {
        "metric": "grid points solved/sec, mixed+refined vs FP64, 512^3; serving p50/p99 latency",
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
            "workload": f"plain FP64 drop-in solve, {n}^3, bc x DD / y ND / z DD, box [0,1]^3, "
                        "RHS U[-1,1] seed 0 (counter generator); value = FP64 form, "
                        f"'mixed' = FP32 + {args.refine} FP64-residual refinement steps on the same RHS",
            "n": n,
            "bc": BC,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "tuning": tuning,
        "kernels_ms": kms,
        "kernels_gbs": gbs,
        "copy_gbs": copy_gbs,
        "stream_copy_gbs": stream_copy,
    }
    if mixed is not None:
        out["mixed"] = mixed
    del f, u
    torch.cuda.empty_cache()
    if args.large > 0:
        ls = max(1, min(args.steps, 5))
        out["large_fp64"], u_large = bench_large(pa, torch, args.large, 0, args.refine, ls)
        progress(f"FP64 {args.large}^3 done")
        out["large_mixed"], u_mixed = bench_large(pa, torch, args.large, 1, args.refine, ls, ref=u_large)
        progress(f"mixed {args.large}^3 done")
        del u_large, u_mixed
        torch.cuda.empty_cache()
    if args.batch_count > 0:
        out["batched"] = bench_batched(pa, args.batch_count, max(1, min(args.steps, 5)), args.refine,
                                       autotune=not args.no_autotune)
        progress("batched form done")
    if args.serve_requests > 0:
        rates = [float(v) for v in args.serve_rates.split(",") if v.strip()]
        out["serving"] = bench_serving_sweep(pa, args.serve_requests, rates)
        progress("serving sweep done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(u64_host, n=n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
