#!/usr/bin/env python3
"""Headline benchmark of the Laguerre-cell path on one MI355X.

Default workload (the line's `value`): the ball-clipped 2D cell pass at N = 1,000,000
diracs uniform in [0,1]^2 (seed 0), weights w_i ~ U[0, 1e-6] shifted by r0^2 = 1e-6,
so each cell is clipped by the disc of radius sqrt(w_i + r0^2); domain the unit square
with density 1.  One step = one full pass on resident HBM inputs: grid build (bounding
box, Morton keys, counting sort, per-cell max weight), cell kernel (power cell, disc
clip, cell area, boundary measures, Hessian rows) and the dirac-order CSR in HBM.
`value` is the FP64 pass (the reference's arithmetic); the mixed FP32-clip pass is
reported beside it with its error against the FP64 oracle.

The same run adds four legs as extra keys:
  c1: 2D Newton, N = 10^4 uniform, w0 = 0, uniform targets, to 1e-6 max relative mass
      error, FP64 and mixed, beside the oracle's own full solve on the host cores;
  c3: 3D Newton, N = 10^6 uniform in [0,1]^3, 64^3 lognormal(0, 0.5) voxel density
      (mass 1), to 1e-6, FP64 and mixed, beside the oracle's cell pass at the converged
      weights timed on a bounded sample of the cells;
  c4: 512 batched 2D Newton solves (N_k log-uniform in [1e3, 1e5], mixed), one solve after
      one warmup (which runs the autotuner), beside the oracle's solves of the first problems;
  c5: the serving loop for a 10 s arrival window at 12 requests/s (70 % 2D / 30 % 3D,
      N log-uniform in [1e3, 2e6]): p50 / p99 time to converged weights.

Other workloads: --workload c3 | c4 | c5 (one line each).
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2] [--no-legs]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Laguerre cells/s (1M diracs 2D/3D); solve s @1e-6 mass err; p99 latency"
PEAK = {"fp64": 78.6, "mixed": 157.3}        # MI355X vector FP64 / FP32, TFLOP/s
# HBM bytes per pass of the cell kernels on the 1M ball workload (all capacity tiers): rocprofv3
# FETCH_SIZE and WRITE_SIZE of k_cells2, each scaled by the same counter on the library's copy
# kernel moving 1 GiB (bench.py --copy-calibration 1024: FETCH reads half the copied bytes,
# WRITE all of them, so FETCH x 2, WRITE x 1)
TRAFFIC = {"mixed": 184.3e6, "fp64": 577.9e6}
# VALU instructions (wavefront instructions, SQ_INSTS_VALU) per pass of the same kernels, and
# the issue rate they are compared with: one wave64 VALU instruction per SIMD every 4 cycles,
# 1024 SIMDs at 2.4 GHz
VALU_INSTS = {"mixed": 4.099e8, "fp64": 4.174e8}
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4
CPU_CORES = min(16, os.cpu_count() or 1)     # the host share of one GPU on the box


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def make_inputs(n, seed=0):
    import sdot_amd
    u = sdot_amd.rng_uniform(seed, 3 * n)
    pos = u[: 2 * n].reshape(n, 2).copy()
    w = 1e-6 * u[2 * n:] + 1e-6
    return pos, w


def run_gpu(pos, w, precision, steps, warmup):
    import sdot_amd
    pd = sdot_amd.PowerDiagram(2, precision, ball=True)
    pd.add_box_shape([0.0, 0.0], [1.0, 1.0])
    pd.set_diracs(pos, w)
    for _ in range(warmup):
        pd.mark_diracs_changed()
        pd.compute()
        pd.csr_device()
    tims = []
    t0 = time.perf_counter()
    for _ in range(steps):
        pd.mark_diracs_changed()
        pd.compute()
        pd.csr_device()
        tims.append(pd.last_timings())
    t1 = time.perf_counter()
    masses = pd.integrals()
    ms_step = (t1 - t0) * 1e3 / steps
    kern = float(np.mean([t["ms_cell_kernel"] for t in tims]))
    flops = tims[-1]["cell_flops"]
    return dict(precision=precision, ms_per_step=ms_step, cells_per_s=len(w) / (ms_step * 1e-3),
                ms_cell_kernel=kern, ms_grid=float(np.mean([t["ms_grid"] for t in tims])),
                ms_rows=float(np.mean([t["ms_rows"] for t in tims])),
                cell_flops=flops, achieved_tflops=flops / (kern * 1e-3) / 1e12,
                overflow_cells=int(tims[-1]["nb_cells_overflow"]), status=pd.last_status), masses


def copy_calibration(mib, reps=3):
    """Plain device-to-device copies of a known size (reads and writes mib MiB each)
    for scaling FETCH_SIZE / WRITE_SIZE counters: the library's copy kernel
    (k_copy16) on buffers of the HIP runtime the library links (one runtime in the
    process)."""
    import ctypes
    import sdot_amd  # noqa: F401  (loads the library and its runtime first)
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    nbytes = mib * (1 << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), nbytes) != 0:
            raise RuntimeError("hipMalloc failed in the copy calibration")
    try:
        if hip.hipMemset(a, 1, nbytes) != 0:
            raise RuntimeError("hipMemset failed in the copy calibration")
        lib = sdot_amd.lib()
        for _ in range(reps):
            if lib.sdot_calibration_copy(b, a, nbytes) != 0:
                raise RuntimeError("copy kernel failed in the copy calibration")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    print(f"# copy calibration: {reps} copies of {mib} MiB", file=sys.stderr, flush=True)


def run_cpu_baseline(pos, w, budget_s=10.0, max_reps=40):
    from oracle import oracle as O
    reps = 0
    t0 = time.perf_counter()
    while True:
        r = O.cells(pos, w, ball=True, nthreads=CPU_CORES)
        O.to_csr(r)
        reps += 1
        if time.perf_counter() - t0 >= budget_s or reps >= max_reps:
            break
    dt = (time.perf_counter() - t0) / reps
    return dict(value=len(w) / dt, unit="cells/s", cores=CPU_CORES, kind="port", cpu_model=cpu_model(),
                sample=f"{reps} full passes of the same 1M-dirac ball workload (grid + cells + CSR) "
                       f"by the C oracle with {CPU_CORES} threads, {dt:.3f} s each"), r["masses"]


def run_c1(reps=3):
    """2D Newton, N = 10^4 (seed 0), w0 = 0, targets 1/N, to 1e-6: device solves (median
    of reps) in both precisions and the oracle's complete solve on the host cores."""
    import sdot_amd
    from sdot_amd import workloads
    from oracle import oracle as O
    n = 10_000
    pos = workloads.uniform_diracs(n, 2, seed=0)
    t = np.full(n, 1.0 / n)
    out = {}
    for prec in ("fp64", "mixed"):
        pd = sdot_amd.PowerDiagram(2, prec)
        pd.add_box_shape([0, 0], [1, 1])
        pd.set_diracs(pos, np.zeros(n))
        runs, stats = [], None
        for _ in range(reps + 1):   # the first solve builds the solver's buffers
            t0 = time.perf_counter()
            st, w, stats = pd.ot_solve(t, np.zeros(n), tol=1e-6)
            runs.append(time.perf_counter() - t0)
        out[prec] = dict(seconds=float(np.median(runs[1:])), status=st, newton_iters=stats["newton_iters"],
                         pcg_iters=stats["pcg_iters_total"], max_rel_mass_err=stats["max_rel_mass_err"],
                         nnz=stats["nnz"])
        out[prec + "_w"] = w
    t0 = time.perf_counter()
    sto, wo, so = O.ot_solve(pos, np.zeros(n), t, tol=1e-6, nthreads=CPU_CORES)
    cpu_s = time.perf_counter() - t0
    d = wo - wo.mean()
    werr = {p: float(np.abs((out[p + "_w"] - out[p + "_w"].mean()) - d).max() / np.abs(d).max())
            for p in ("fp64", "mixed")}
    res = {"workload": "2D Newton OT solve, N=10000 uniform in [0,1]^2 (seed 0), w0=0, targets 1/N, unit box, "
                       "stop at max relative mass error 1e-6",
           "fp64": out["fp64"], "mixed": out["mixed"],
           "weights_rel_diff_vs_oracle": werr,
           "cpu_baseline": {"value": cpu_s, "unit": "s", "cores": CPU_CORES, "kind": "port",
                            "sample": f"the oracle's complete Newton solve of the same problem (Jacobi-PCG), "
                                      f"{so.newton_iters} iterations, status {sto}"},
           "speedup_vs_cpu": {p: cpu_s / out[p]["seconds"] for p in ("fp64", "mixed")}}
    print(f"# c1: {res}", file=sys.stderr, flush=True)
    return res


def run_c3(n, img_size, steps, precisions=("fp64", "mixed")):
    """3D Newton solve: n diracs uniform in [0,1]^3, lognormal(0, 0.5) voxel density on an
    img_size^3 grid normalised to mass 1, w0 = 0, uniform targets, stop at 1e-6 relative mass error."""
    import sdot_amd
    from sdot_amd import workloads
    pos = workloads.uniform_diracs(n, 3, seed=0)
    img = workloads.lognormal_image(img_size, 3, seed=1)
    t = np.full(n, img.mean() / n)
    out = {}
    for prec in precisions:
        pd = sdot_amd.PowerDiagram(3, prec)
        pd.set_image_density([0, 0, 0], [1, 1, 1], img)
        pd.set_diracs(pos, np.zeros(n))
        runs = []
        for _ in range(steps):
            t0 = time.perf_counter()
            st, w, stats = pd.ot_solve(t, np.zeros(n), tol=1e-6)
            runs.append(time.perf_counter() - t0)
        stats["seconds"] = float(np.median(runs))
        stats["status"] = st
        stats["cells_per_s_pass"] = n * stats["newton_iters"] / stats["seconds_cells"] if stats["seconds_cells"] else None
        stats["ms_pcg_iteration"] = 1e3 * stats["seconds_solve"] / max(1, stats["pcg_iters_total"])
        # one more pass at the converged weights: the wave kernel against the vector peak
        pd.set_weights(w)
        pd.integrals()
        tm = pd.last_timings()
        ach = tm["cell_flops"] / (tm["ms_cell_kernel"] * 1e-3) / 1e12
        stats["roofline_pass"] = {"bound": "valu", "kernel": "k_cells3_wave", "flops_per_launch": tm["cell_flops"],
                                  "kernel_ms": tm["ms_cell_kernel"], "achieved": ach, "peak": PEAK[prec],
                                  "unit": "TFLOP/s", "frac": ach / PEAK[prec]}
        out[prec] = stats
        out[prec + "_weights"] = w
        print(f"# c3 {prec}: {stats}", file=sys.stderr, flush=True)
    return out, pos, img


def c3_cpu_baseline(pos, w, img, budget_s=10.0):
    """The oracle's cell pass (voxel pieces, Hessian rows) at the converged weights, on the
    first S cells in dirac order (a spatially uniform sample), S sized to about budget_s."""
    from oracle import oracle as O
    n = len(pos)
    s0 = min(n, 2000)
    dt0, _, _ = O.cells_sample_seconds(pos, w, 0, s0, img=img, nthreads=CPU_CORES)
    s = int(min(n, max(s0, s0 * budget_s / max(dt0, 1e-6))))
    dt, st, _ = O.cells_sample_seconds(pos, w, 0, s, img=img, nthreads=CPU_CORES)
    return dict(value=s / dt, unit="cells/s (3D image pass at the converged weights)", cores=CPU_CORES, kind="port",
                cpu_model=cpu_model(),
                sample=f"oracle cell pass (voxel pieces, Hessian rows) on cells [0, {s}) of the {n} at the "
                       f"converged FP64 weights: {dt:.2f} s with {CPU_CORES} threads")


def c3_leg(n=1_000_000, img=64):
    res, pos, image = run_c3(n, img, 1)
    cpu = c3_cpu_baseline(pos, res["fp64_weights"], image)
    f, m = res["fp64"], res["mixed"]
    # a lower bound of the CPU solve: its Newton passes alone (every pass a full pass at the
    # measured rate), without its linear solves
    cpu_passes = f["newton_iters"] * n / cpu["value"]
    out = {"workload": f"3D Newton OT solve, N={n} uniform in [0,1]^3, {img}^3 lognormal(0,0.5) voxel density "
                       "(mass 1), w0=0, uniform targets, stop at max relative mass error 1e-6",
           "fp64": {k: f[k] for k in ("seconds", "status", "newton_iters", "pcg_iters_total", "backtracks",
                                      "max_rel_mass_err", "seconds_cells", "seconds_solve", "cells_per_s_pass",
                                      "ms_pcg_iteration", "roofline_pass")},
           "mixed": {k: m[k] for k in ("seconds", "status", "newton_iters", "pcg_iters_total", "backtracks",
                                       "max_rel_mass_err", "seconds_cells", "seconds_solve", "cells_per_s_pass",
                                       "ms_pcg_iteration", "roofline_pass")},
           "mixed_speedup_vs_fp64": f["seconds"] / m["seconds"],
           "cpu_baseline": cpu,
           "gpu_pass_vs_cpu_pass": {p: res[p]["cells_per_s_pass"] / cpu["value"] for p in ("fp64", "mixed")},
           "cpu_solve_lower_bound_s": cpu_passes,
           "solve_speedup_lower_bound": {p: cpu_passes / res[p]["seconds"] for p in ("fp64", "mixed")}}
    print(f"# c3 leg: {out}", file=sys.stderr, flush=True)
    return out


def main_c3(args):
    precs = ("fp64", "mixed") if args.precision == "both" else (args.precision,)
    res, pos, img = run_c3(args.n, args.img, args.steps, precs)
    best = "fp64" if "fp64" in precs else precs[0]
    cpu = None if args.no_cpu_baseline else c3_cpu_baseline(pos, res[best + "_weights"], img)
    line = {
        "metric": "solve s @1e-6 mass err (3D Newton, 1M diracs, 64^3 lognormal density)",
        "value": res[best]["seconds"], "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": res[best]["seconds"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if best == "fp64" else "f32-clip/f64-solve",
        "data": "synthetic (xoshiro256** seed 0 positions, seed 1 density)",
        "config": {"workload": f"3D Newton OT solve, N={args.n} uniform in [0,1]^3, {args.img}^3 lognormal(0,0.5) "
                               "voxel density (mass 1), w0=0, uniform targets, stop at max relative mass error 1e-6",
                   "n_diracs": args.n, "dim": 3, "precision": best},
        "solves": {p: {k: v for k, v in res[p].items()} for p in precs},
        "mixed_speedup_vs_fp64": (res["fp64"]["seconds"] / res["mixed"]["seconds"]) if len(precs) == 2 else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def run_c4(K, steps, warmup, modes, use_graph=True, cpu_budget_s=10.0):
    """512 independent 2D Newton solves in one batch (mixed precision, per-problem
    stop at 1e-6 relative mass error).  Problem k: N_k = floor(exp(U[ln 1e3, ln 1e5]))
    (U from seed 0), positions uniform from seed k+1, w0 = 0, targets 1/N_k, unit box.
    One step = one sdot_batch_solve of all problems (grids, cells, PCG, Newton control).
    cpu_budget_s = 0: no CPU baseline."""
    import sdot_amd
    from sdot_amd import workloads
    from oracle import oracle as O
    sizes = workloads.batch_sizes(K, 1e3, 1e5, seed=0)
    pos = [workloads.uniform_diracs(int(n), 2, seed=k + 1) for k, n in enumerate(sizes)]
    b = sdot_amd.Batch("mixed", diracs_per_cell=0.0, use_graph=use_graph)
    b.set_problems(pos)
    timed = {}
    for mode in modes:   # the first mode's warmup also runs the autotuner
        b.set_schedule(device_newton=mode == "device")
        for _ in range(warmup):
            r = b.solve(tol=1e-6)
            print(f"# c4 {mode} warmup: status {r[0]} stats {r[5]}", file=sys.stderr, flush=True)
        runs = []
        for _ in range(steps):
            t0 = time.perf_counter()
            res = b.solve(tol=1e-6)
            runs.append(time.perf_counter() - t0)
            print(f"# c4 {mode} step {len(runs)}: {runs[-1]:.3f} s status {res[0]} stats {res[5]}", file=sys.stderr,
                  flush=True)
        timed[mode] = (float(np.sum(runs) / len(runs)), res)
    sec, (st, ws, ms, pst, its, stats) = timed[modes[0]]
    # CPU: the oracle's own solve on the first problems, about cpu_budget_s, scaled by dirac count
    cpu = None
    if cpu_budget_s > 0:
        cores = CPU_CORES
        t0 = time.perf_counter()
        ns = 0
        k = 0
        while k < K and (time.perf_counter() - t0 < cpu_budget_s or k == 0):
            n = len(pos[k])
            O.ot_solve(pos[k], np.zeros(n), np.full(n, 1.0 / n), tol=1e-6, nthreads=cores)
            ns += n
            k += 1
        cpu_s = time.perf_counter() - t0
        cpu_est = cpu_s * float(sizes.sum()) / ns
        cpu = {"value": cpu_est, "unit": "s (estimate)", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"oracle Newton solves of the first {k} problems ({ns} diracs) took "
                         f"{cpu_s:.2f} s; scaled by total diracs / sampled diracs"}
    err = max(float(np.abs(ms[i] * len(ms[i]) - 1).max()) for i in range(K))
    return dict(seconds=sec, modes={m: timed[m][0] for m in modes}, its=its, stats=stats, err=err, cpu=cpu,
                total_diracs=int(sizes.sum()))


def c4_leg(K=512):
    r = run_c4(K, 1, 1, ["device"], cpu_budget_s=5.0)
    its = r["its"]
    return {"workload": f"{K} batched 2D Newton OT solves, N_k log-uniform in [1e3,1e5] (sum {r['total_diracs']}), "
                        "mixed, per-problem stop at max relative mass error 1e-6, Newton driven from the device",
            "seconds": r["seconds"], "max_rel_mass_err": r["err"],
            "newton_iters_per_problem": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
            "pcg_iters_total": r["stats"]["pcg_iters_total"], "cpu_baseline": r["cpu"],
            "speedup_vs_cpu_estimate": r["cpu"]["value"] / r["seconds"] if r["cpu"] else None}


def main_c4(args):
    K = args.problems
    modes = ["device", "host"] if args.newton == "both" else [args.newton]
    r = run_c4(K, args.steps, args.warmup, modes, use_graph=not args.no_graph,
               cpu_budget_s=0.0 if args.no_cpu_baseline else 10.0)
    sec, its, stats, err, cpu = r["seconds"], r["its"], r["stats"], r["err"], r["cpu"]
    line = {
        "metric": "solve s @1e-6 mass err (512 batched 2D Newton problems)",
        "value": sec, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-clip/f64-solve", "data": "synthetic (xoshiro256** seeds 0..512)",
        "config": {"workload": f"{K} independent 2D OT Newton solves, N_k log-uniform in [1e3,1e5] "
                               f"(sum {r['total_diracs']}), unit box, w0=0, targets 1/N_k, per-problem stop at max "
                               "relative mass error 1e-6, " + (
                                   "Newton driven from the device (one captured graph per iteration, replayed)"
                                   if modes[0] == "device" else "host-driven Newton iterations, PCG from a captured "
                                   "hipGraph"),
                   "problems": K, "total_diracs": r["total_diracs"], "precision": "mixed", "newton": modes[0]},
        "newton_modes_s": r["modes"],
        "newton_iters_per_problem": {"min": int(its.min()), "median": float(np.median(its)), "max": int(its.max())},
        "ms_pcg_iteration_graph": stats["ms_pcg_iteration"],
        "max_rel_mass_err": err,
        "stats": stats,
        "cpu_baseline": cpu,
        "speedup_vs_cpu_estimate": cpu["value"] / sec if cpu else None,
    }
    print(json.dumps(line))


def c5_schedule(rate, duration, seed=7):
    """Poisson arrivals (rate/s over `duration` s), 70 % 2D / 30 % 3D, N = floor(exp(U[ln 1e3, ln 2e6]))."""
    from sdot_amd import workloads
    return workloads.serving_schedule(rate, duration, seed)


def c5_run(rate, duration, deadline, precision, seed=7, label=""):
    """Open-loop load: submit each request at its arrival time, collect until `deadline` s."""
    import threading
    import sdot_amd
    from sdot_amd import workloads
    t_arr, dims, sizes = c5_schedule(rate, duration, seed)
    srv = sdot_amd.Server(precision, diracs_per_cell=2.5, workers_3d=2, tol=1e-6, hbm_budget_bytes=64e9)
    ids = [None] * len(t_arr)
    done_submit = threading.Event()
    t0 = time.perf_counter()

    def submitter():
        for k in range(len(t_arr)):
            dt = t_arr[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            if time.perf_counter() - t0 > deadline:
                break
            pos = workloads.uniform_diracs(int(sizes[k]), int(dims[k]), seed=10_000 + k)
            ids[k] = srv.submit(pos)
        done_submit.set()

    th = threading.Thread(target=submitter, daemon=True)
    th.start()
    # collect every finished request, in whatever order they finish: each sweep polls
    # all outstanding ids (wait with a zero timeout returns at once when one is pending)
    res = {}
    outstanding = set()
    seen = 0
    last = time.perf_counter()
    while True:
        while seen < len(ids) and ids[seen] is not None:
            outstanding.add(seen)
            seen += 1
        for k in list(outstanding):
            r = srv.wait(ids[k], timeout=0.0)
            if r is not None:
                res[k] = r[1]
                outstanding.discard(k)
        if done_submit.is_set() and seen >= sum(1 for i in ids if i is not None) and not outstanding:
            break
        if time.perf_counter() - t0 > deadline:
            break
        if time.perf_counter() - last > 20:
            print(f"# c5{label}: {len(res)} done, {len(outstanding)} outstanding, stats {srv.stats()}",
                  file=sys.stderr, flush=True)
            last = time.perf_counter()
        time.sleep(0.005)
    wall = time.perf_counter() - t0
    th.join(timeout=1.0)
    st = srv.stats()
    srv.close()
    lat = np.array([r["t_done"] - r["t_submit"] for r in res.values()])
    ok = sum(1 for r in res.values() if r["status"] == 0)
    submitted = sum(1 for i in ids if i is not None)
    cells = st["cells_evaluated"]
    out = {"rate_per_s": rate, "arrival_window_s": duration, "offered": int(len(t_arr)), "submitted": submitted,
           "completed": len(res), "converged": ok, "wall_s": wall,
           "throughput_req_per_s": len(res) / wall, "cells_per_s": cells / wall,
           "p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
           "p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
           "max_s": float(lat.max()) if len(lat) else None,
           "backlog_at_end": submitted - len(res), "server": st,
           "offered_diracs": int(sizes.sum()), "completed_diracs": int(sum(sizes[i] for i in res))}
    print(f"# c5{label}: {out}", file=sys.stderr, flush=True)
    return out


def c5_leg(rate=12.0, window=10.0):
    """A bounded serving run at a rate the device sustains (every request served)."""
    r = c5_run(rate, window, window + 120.0, "mixed", seed=8, label=" leg")
    keys = ("rate_per_s", "arrival_window_s", "offered", "completed", "converged", "throughput_req_per_s",
            "cells_per_s", "p50_s", "p99_s", "max_s", "backlog_at_end")
    out = {k: r[k] for k in keys}
    out["workload"] = ("serving loop: Poisson arrivals, 70 % 2D / 30 % 3D, N = floor(exp(U[ln 1e3, ln 2e6])), unit box, "
                       "uniform targets, w0 = 0, stop at max relative mass error 1e-6, mixed; latency = time to "
                       "converged weights")
    return out


def main_c5(args):
    """Serving loop.  Phase 1: the offered load (Poisson 200/s, 70 % 2D / 30 % 3D,
    N log-uniform in [1e3, 2e6]) for a bounded arrival window; the queue grows
    without bound at this rate, so completed requests, throughput and backlog
    are reported.  Phase 2: a sustainable rate (half the phase-1 throughput),
    every request served; p50/p99 time-to-converged-weights."""
    over = c5_run(args.rate, args.window, args.window + args.drain, "mixed", label=" offered")
    sus_rate = max(0.2, 0.5 * over["throughput_req_per_s"])
    sus = c5_run(sus_rate, args.sustain, args.sustain + 240.0, "mixed", seed=8, label=" sustainable")
    line = {
        "metric": "p99 latency s (serving loop, time to converged weights @1e-6)",
        "value": sus["p99_s"], "unit": "s", "n_gpus": 1, "steps": sus["completed"], "warmup": 0,
        "ms_per_step": (sus["p50_s"] or 0) * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-clip/f64-solve", "data": "synthetic (xoshiro256** seeds 7/8, 10000+k)",
        "config": {"workload": "serving loop: Poisson arrivals, 70 % 2D / 30 % 3D, N = floor(exp(U[ln 1e3, ln 2e6])), "
                               "unit box, uniform targets, w0 = 0, stop at max relative mass error 1e-6; "
                               f"value = p99 at a sustainable {sus_rate:.2f} req/s",
                   "precision": "mixed"},
        "sustainable": sus, "offered_load": over,
        "cpu_baseline": None,
    }
    print(json.dumps(line))




def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5"], default="c2")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--window", type=float, default=10.0)
    ap.add_argument("--drain", type=float, default=50.0)
    ap.add_argument("--sustain", type=float, default=30.0)
    ap.add_argument("--problems", type=int, default=512)
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--img", type=int, default=64)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="c2 only: skip the C1 / C3 / C4 / C5 legs")
    ap.add_argument("--no-graph", action="store_true", help="c4 only: launch the kernels one by one")
    ap.add_argument("--newton", choices=["device", "host", "both"], default="both",
                    help="c4 only: device-driven Newton (the line's value), host-driven, or both")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run device-to-device copies of MIB MiB (known HBM bytes) so hardware "
                         "counters of the cell kernel can be scaled by the copy's counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only (the cell pass of one GPU's share); --gpus {args.gpus} "
              "is not run: multi-GPU sharding is a design (ghost-layer exchange), not part of this build.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.workload == "c3":
        args.steps = min(args.steps, 3)
        return main_c3(args)
    if args.workload == "c4":
        args.steps = min(args.steps, 5)
        args.warmup = max(1, min(args.warmup, 2))
        return main_c4(args)
    if args.workload == "c5":
        return main_c5(args)

    pos, w = make_inputs(args.n)
    res = {}
    masses = {}
    for prec in ("fp64", "mixed"):
        res[prec], masses[prec] = run_gpu(pos, w, prec, args.steps, args.warmup)
        print(f"# {prec}: {res[prec]}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu, ref = run_cpu_baseline(pos, w)
        scale = 1.0 / args.n
        for prec in res:
            err = np.abs(masses[prec] - ref)
            res[prec]["max_mass_err_vs_oracle"] = float((err / (np.abs(ref) + 1e-9 * scale)).max())
            res[prec]["max_mass_err_x_N"] = float(err.max() / scale)
    legs = {}
    if not args.no_legs:
        legs["c1"] = run_c1()
        legs["c3"] = c3_leg()
        legs["c4"] = c4_leg()
        legs["c5"] = c5_leg()

    b = res["fp64"]
    achieved = b["achieved_tflops"]
    line = {
        "metric": METRIC,
        "value": b["cells_per_s"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (xoshiro256** seed 0)",
        "config": {"workload": "2D ball-clipped cell pass, N=1e6 uniform in [0,1]^2, w~U[0,1e-6]+r0^2, r0=1e-3, "
                               "unit-square domain; grid + cells + Hessian CSR per step",
                   "n_diracs": args.n, "dim": 2, "precision": "fp64"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK["fp64"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK["fp64"], "traffic": TRAFFIC["fp64"] if args.n == 1_000_000 else None,
                     "kernel": "k_cells2 (FP64 capacity tiers)", "flops_per_launch": b["cell_flops"],
                     "kernel_ms": b["ms_cell_kernel"],
                     "valu_issue_frac": VALU_INSTS["fp64"] / (b["ms_cell_kernel"] * 1e-3 * VALU_ISSUE_PER_S)
                     if args.n == 1_000_000 else None},
        "cpu_baseline": cpu,
        "mixed": {"cells_per_s": res["mixed"]["cells_per_s"],
                  "speedup_vs_fp64": res["mixed"]["cells_per_s"] / b["cells_per_s"],
                  "dtype": "f32-clip/f64-integrate",
                  "max_mass_err_vs_oracle": res["mixed"].get("max_mass_err_vs_oracle"),
                  "roofline": {"achieved": res["mixed"]["achieved_tflops"], "peak": PEAK["mixed"],
                               "frac": res["mixed"]["achieved_tflops"] / PEAK["mixed"],
                               "traffic": TRAFFIC["mixed"] if args.n == 1_000_000 else None,
                               "valu_issue_frac": VALU_INSTS["mixed"] / (res["mixed"]["ms_cell_kernel"] * 1e-3 *
                                                                         VALU_ISSUE_PER_S)
                               if args.n == 1_000_000 else None}},
        "precisions": res,
        "host": {"cpu_model": cpu_model(), "cores_used": CPU_CORES},
    }
    line.update(legs)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
