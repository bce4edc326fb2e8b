#!/usr/bin/env python3
"""Throughput of the spherical Vlasov-Poisson spline step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: the UDF uniform-density sphere (M=1, R0=2, alpha=0.5, smoothing
0.5, G=1) on its default grid, 201 log-r nodes (0.01..25) x 201 linear v_r
nodes (-2..2) x 200 quadratic j cells (0..1.6), dt=0.005, FP64: 8,080,200
phase-space points.  A step is one full makeOneStep (moment, radial scan,
kick spline + evaluation, drift spline + evaluation, delay lines), with f
resident in HBM.  value = points x K / wall seconds of K steps.

Prints one JSON line with, besides the throughput, a roofline object: for
the dominant kernel, its algorithmic bytes per launch over its average
duration, measured with HIP events on the solver's stream placed just before
and just after each of its launches (one kernel timed per pass, the other
kernels undisturbed), the per-step time of every kernel measured the same
way, and the whole step's algorithmic bytes over ms_per_step; a
cpu_baseline object: the reference executable, built from its own sources
into oracle/_ref, timed on this host's cores (all of this process's CPUs, at
most 16, and one) on a bounded sample of the same workload (or the oracle,
as a port, if the reference was not built), with nproc and the lscpu
summary; and the other workloads' legs: the PLDF 129x129x64 adjoint/forward
cost ratio, the 32-run ensemble speedup, and the Hernquist 513x513x256
forward throughput and adjoint/forward cost ratio.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

WORKLOAD = {
    "init.type": "UDF", "init.M": 1, "init.R0": 2, "init.alpha": 0.5,
    "init.smoothingRadius": 0.5, "solver.G": 1,
    "grid.Pmin": 0.01, "grid.Pmax": 25, "grid.Pres": 200, "grid.Pscale": "logarithmic",
    "grid.Umin": [-2, 0], "grid.Umax": [2, 1.6], "grid.Ures": [200, 200],
    "grid.Uscale": ["linear", "quadratic"], "solver.dt": 0.005,
}


def algorithmic_bytes(name, nR, nU, nJ):
    """Compulsory HBM bytes of one launch of a kernel of the step: every
    operand array read once and written once; static foot and delay-line
    tables excluded."""
    n = nR * nU * nJ
    coef = (nR + 2) * (nU + 2) * nJ
    rows = (nR + 2) * nU * nJ
    cols = nR * (nU + 2) * nJ            # the kick's column splines
    table = {
        "moment": 8 * n,                 # read f
        "spline_r": 8 * n + 8 * rows,    # read f, write the r-solved rows
        "spline_v": 8 * rows + 8 * coef,  # read them, write all coefficients
        "kick": 16 * n + 16 * cols,      # column splines of f, then the kick: f -> d -> f
        "kick_line": 8 * n + 8 * n,      # fused column splines + kick: read f, write f
        "kick_solve": 8 * n + 8 * cols,  # column splines and moment partials: read f, write d
        "kick_eval": 8 * cols + 8 * n,   # the kick from d: read d, write f
        "kick_spline_r": 8 * cols + 8 * rows,  # the kicked rows' r-solve: read d, write rows
        "drift": 8 * coef + 8 * n,       # read coefficients, write f
    }
    return table.get(name)


# The kernels of one step of the workload, as vls_set_timing_kernel names them
# ("moment" covers the moment partials, the chunk sums and the scan;
# "kick_solve" the kick's column splines with the moment; "kick_spline_r" the
# kick's evaluation folded into the drift's r-solve).
STEP_KERNELS = ("moment", "kick_solve", "kick_eval", "kick_spline_r", "kick_line", "kick",
                "spline_r", "spline_v", "drift")
STEP_BYTES_PER_POINT = 72

# The adjoint's workload: PLDF power-law sphere (n=1, alpha=1, eta=0.1, M=1,
# smoothing 0.5) on 129 log-r x 129 v x 64 quadratic-j nodes, dt=0.005,
# 2000 steps, checkpoints every 50 steps.
ADJ_WORKLOAD = {
    "init.type": "PLDF", "init.n": 1, "init.alpha": 1, "init.eta": 0.1, "init.M": 1,
    "init.smoothingRadius": 0.5, "grid.Pres": 128, "grid.Ures": [128, 64], "grid.Pmax": 30,
    "grid.Umin": [-1.5, 0], "grid.Umax": [1.5, 1.6], "solver.dt": 0.005,
}
ADJ_STEPS, ADJ_CK = 2000, 50


def adjoint_ratio(V):
    """wall(gradient of a mass-profile loss: recorded forward run with
    checkpoints + reverse sweep recomputing each segment) / wall(forward).
    ratio: the first gradient, which also builds the drift's transposed foot
    table on the host and allocates the checkpoints; ratio_steady: a second
    one."""
    s = V.Solver(ADJ_WORKLOAD)
    ctx = s.context
    f0 = s.f()
    head0, vals0 = ctx.get_chains()
    ctx.step(20)  # warm-up
    ctx.synchronize()
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, vals0)
    t0 = time.perf_counter()
    ctx.step(ADJ_STEPS)
    ctx.synchronize()
    fwd = time.perf_counter() - t0
    grad = _gradient_wall(ctx, f0, head0, vals0, ADJ_STEPS, ADJ_CK)
    grad2 = _gradient_wall(ctx, f0, head0, vals0, ADJ_STEPS, ADJ_CK)
    return {"workload": "PLDF n=1 eta=0.1, 129 log-r x 129 v x 64 quadratic-j, dt 0.005",
            "steps": ADJ_STEPS, "ck_every": ADJ_CK, "forward_s": round(fwd, 4),
            "gradient_s": round(grad, 4), "ratio": round(grad / fwd, 3),
            "gradient_s_steady": round(grad2, 4), "ratio_steady": round(grad2 / fwd, 3),
            "loss": "sum_r (M(<r) - M_target)^2 of the final state"}


# The large adjoint workload: the Hernquist sphere (M=1, R0=1) on its default
# ranges, 513 log-r x 513 v x 256 log-j nodes (67,371,264 points), dt=0.0025,
# 4000 steps with checkpoints every 63 steps (about sqrt(4000)): about 1.5
# minutes of GPU time and 80 GB of HBM (--no-hernquist skips it).
HQ_WORKLOAD = {"init.type": "Hernquist", "init.M": 1, "init.R0": 1, "grid.Pres": 512,
               "grid.Ures": [512, 256], "solver.dt": 0.0025}
HQ_STEPS, HQ_CK = 4000, 63


def _gradient_wall(ctx, f0, head0, vals0, steps, ck):
    import numpy as np

    ctx.upload_f(0, f0)
    ctx.put_chains(head0, vals0)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.record(steps, ck)
    M = ctx.mass_profile()
    g0, _ = ctx.adjoint(dL_dM=2 * (M - 1.01 * M))
    wall = time.perf_counter() - t0
    if not np.isfinite(g0).all():
        raise RuntimeError("non-finite gradient")
    return wall


def adjoint_ratio_hernquist(V, steps=HQ_STEPS, ck=HQ_CK):
    """The adjoint/forward wall ratio on the Hernquist workload, first and
    steady gradient as in adjoint_ratio."""
    s = V.Solver(HQ_WORKLOAD)
    try:
        return _hernquist_legs(s, steps, ck)
    finally:
        s.close()  # its kept steps hold most of the HBM


def _hernquist_legs(s, steps, ck):
    ctx = s.context
    f0 = s.f()
    head0, vals0 = ctx.get_chains()
    ctx.step(3)
    ctx.upload_f(0, f0)
    ctx.put_chains(head0, vals0)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    fwd = time.perf_counter() - t0
    first = _gradient_wall(ctx, f0, head0, vals0, steps, ck)
    second = _gradient_wall(ctx, f0, head0, vals0, steps, ck)
    npts = s.nR * s.nU * s.nJ
    return {"workload": "Hernquist M=1 R0=1, 513 log-r x 513 v x 256 log-j, dt 0.0025",
            "points": npts, "steps": steps, "ck_every": ck, "forward_s": round(fwd, 3),
            "forward_point_updates_per_s": npts * steps / fwd,
            "gradient_s": round(first, 3), "ratio": round(first / fwd, 3),
            "gradient_s_steady": round(second, 3), "ratio_steady": round(second / fwd, 3),
            "loss": "sum_r (M(<r) - M_target)^2 of the final state"}


# The batched adjoint: 8 Hernquist runs (M=1, R0=1, default ranges) on
# uneven grids, each axis drawn from 128..384 cells (seeded), dt=0.0025,
# 200 steps with checkpoints every 14 (about sqrt(200)); each member's f0
# perturbed by a seeded multiplicative 1e-3, its loss the mismatch of its
# final mass profile against the unperturbed run's.
HQB_STEPS, HQB_CK, HQB_SEED = 200, 14, 7


def hernquist_batch_cfgs(n=8, seed=HQB_SEED):
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        pr, pu, pj = (int(x) for x in rng.integers(128, 385, size=3))
        out.append(dict(HQ_WORKLOAD, **{"grid.Pres": pr, "grid.Ures": [pu, pj]}))
    return out


def adjoint_batch_ratio(V, steps=HQB_STEPS, ck=HQB_CK, seed=HQB_SEED):
    """wall(gradients of the 8-run batch: vls_record_batch + vls_adjoint_batch)
    / wall(the batch's forward run), every run on a stream of its own."""
    import numpy as np

    cfgs = hernquist_batch_cfgs()
    e = V.Ensemble(cfgs)
    try:
        return _batch_legs(e, steps, ck, seed)
    finally:
        e.close()  # its kept steps hold most of the HBM


def _batch_legs(e, steps, ck, seed):
    import numpy as np

    ctx = e.context
    f0 = [e.f(k) for k in range(e.n)]
    ch0 = [ctx.get_chains(k) for k in range(e.n)]
    e.step(steps)  # the unperturbed runs give the target profiles
    Mt = [ctx.mass_profile(k) for k in range(e.n)]
    rng = np.random.default_rng(seed)
    fp = [f * (1 + 1e-3 * rng.standard_normal(f.shape)) for f in f0]

    def reset():
        for k in range(e.n):
            ctx.upload_f(k, fp[k])
            ctx.put_chains(ch0[k][0], ch0[k][1], run=k)
        ctx.synchronize()

    reset()
    e.step(2)  # warm-up (graph capture)
    reset()
    t0 = time.perf_counter()
    e.step(steps)
    ctx.synchronize()
    fwd = time.perf_counter() - t0
    walls, parts = [], []
    for _ in range(2):
        reset()
        t0 = time.perf_counter()
        ctx.record_batch(steps, ck)
        t1 = time.perf_counter()
        dL = [2 * (ctx.mass_profile(k) - Mt[k]) for k in range(e.n)]
        grads = ctx.adjoint_batch(dL_dM=dL)
        walls.append(time.perf_counter() - t0)
        parts.append((round(t1 - t0, 3), round(walls[-1] - (t1 - t0), 3)))
        if not all(np.isfinite(g[0]).all() for g in grads):
            raise RuntimeError("non-finite gradient")
    pts = sum(f.size for f in f0)
    return {"workload": "8 Hernquist runs M=1 R0=1, axes drawn from 128..384 cells (seed 7), "
                        "f0 x (1 + 1e-3 N(0,1)), dt 0.0025", "points": pts, "steps": steps,
            "ck_every": ck, "grids": [list(f.shape[::-1]) for f in f0],
            "forward_s": round(fwd, 3), "gradient_s": round(walls[0], 3),
            "ratio": round(walls[0] / fwd, 3), "gradient_s_steady": round(walls[1], 3),
            "ratio_steady": round(walls[1] / fwd, 3), "record_reverse_s": parts,
            "loss": "sum_k sum_r (M_k(<r) - M_k,unperturbed(<r))^2 of the final states"}


def ensemble_cfgs(n=32, seed=2024):
    """32 UDF runs: alpha ~ U[0.1, 1], R0 ~ U[1.5, 3], each axis drawn from
    {64, 96, 128, 192, 256} cells (j at most 128)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    sizes = [64, 96, 128, 192, 256]
    out = []
    for _ in range(n):
        pr, pu = int(rng.choice(sizes)), int(rng.choice(sizes))
        pj = int(rng.choice([s for s in sizes if s <= 128]))
        out.append(dict(WORKLOAD, **{"init.alpha": float(rng.uniform(0.1, 1.0)),
                                     "init.R0": float(rng.uniform(1.5, 3.0)),
                                     "grid.Pres": pr, "grid.Ures": [pu, pj]}))
    return out


ENS_STEPS = 500


def ensemble_speedup(V):
    """Throughput of the 32-run uneven ensemble stepped as one graph launch a
    step (runs concurrent) over the same runs stepped one after the other."""
    cfgs = ensemble_cfgs()
    e = V.Ensemble(cfgs)
    pts = sum(e.context.info(k)["nR"] * e.context.info(k)["nU"] * e.context.info(k)["nJ"]
              for k in range(e.n))

    def timed(mode):
        e.set_graph(mode)
        e.step(5)
        e.context.synchronize()
        t0 = time.perf_counter()
        e.step(ENS_STEPS)
        e.context.synchronize()
        return time.perf_counter() - t0

    # the default schedules, then the tuned ones (vls_autotune, timed outside)
    seq, graph = timed(False), timed(True)
    t0 = time.perf_counter()
    schedules = e.autotune()
    tune_s = time.perf_counter() - t0
    seq_t, graph_t = timed(False), timed(True)
    return {"workload": "32 UDF runs, alpha~U[0.1,1], R0~U[1.5,3], axes from {64..256} cells "
                        "(j <= 128), seed 2024", "steps": ENS_STEPS, "points": pts,
            "sequential_s": round(seq, 4), "graph_s": round(graph, 4),
            "speedup": round(seq / graph, 3),
            "tuned_sequential_s": round(seq_t, 4), "tuned_graph_s": round(graph_t, 4),
            "tuned_speedup": round(seq / graph_t, 3),
            "point_updates_per_s": pts * ENS_STEPS / min(graph, graph_t),
            "autotune_s": round(tune_s, 2), "schedules": schedules}


def cpu_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    at most 16 (the CPU share of one GPU on a shared box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_description():
    """nproc and the lscpu summary of the host that runs the baseline."""
    out = {"nproc": None, "affinity": None, "model": None, "sockets": None,
           "cores_per_socket": None, "threads_per_core": None}
    try:
        out["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout)
    except (OSError, ValueError):
        pass
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core"}
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                out[keys[k.strip()]] = int(v) if v.isdigit() else v
    except OSError:
        pass
    return out


def _reference_rate(steps, threads):
    from oracle import oracle as O

    cfg = dict(WORKLOAD)
    cfg.update({"solver.Tmax": 0.005 * steps, "solver.statsEvery": -1,
                "solver.snapshotEvery": -1, "solver.restartEvery": -1})
    with tempfile.TemporaryDirectory() as wd:
        env_bind = os.environ.get("OMP_PROC_BIND")
        os.environ["OMP_PROC_BIND"] = "close"
        try:
            out = O.run_reference(cfg, wd, threads=threads, timeout=900)
        finally:
            if env_bind is None:
                os.environ.pop("OMP_PROC_BIND", None)
            else:
                os.environ["OMP_PROC_BIND"] = env_bind
    m = re.search(r"Computed (\d+) timesteps in ([0-9.eE+-]+)s", out)
    if not m:
        raise RuntimeError("cannot parse the reference's timing line")
    return int(m.group(1)), float(m.group(2))


def cpu_baseline(steps: int, steps_1core: int):
    """Reference executable (or the oracle) on a bounded sample: the same
    grid and initial conditions, stats/snapshots/restarts off, on all the
    host threads this process may use (at most 16) and on one."""
    from oracle import oracle as O

    threads = cpu_threads()
    npts = 201 * 201 * 200
    host = host_description()
    if O.REF_EXE.exists():
        nsteps, secs = _reference_rate(steps, threads)
        n1, s1 = _reference_rate(steps_1core, 1)
        return {"value": npts * nsteps / secs, "unit": "point-updates/s", "cores": threads,
                "kind": "reference", "value_1core": npts * n1 / s1, "host": host,
                "sample": f"reference vlasolve built from its sources (clang -O3, OpenMP, "
                          f"OMP_PROC_BIND=close), same 201x201x200 UDF grid: {nsteps} steps in "
                          f"{secs:.2f} s at OMP_NUM_THREADS={threads}, {n1} steps in {s1:.2f} s "
                          f"at 1 thread (its own step-loop timer)"}
    os.environ["OMP_NUM_THREADS"] = str(threads)
    orc = O.Oracle(WORKLOAD, nthreads=threads)
    orc.step(1)
    t0 = time.perf_counter()
    orc.step(steps)
    secs = time.perf_counter() - t0
    return {"value": npts * steps / secs, "unit": "point-updates/s", "cores": threads,
            "kind": "port", "host": host,
            "sample": f"CPU oracle (C restatement, OpenMP), same 201x201x200 UDF grid, "
                      f"{steps} steps in {secs:.2f} s, {threads} threads"}


def traffic_for(kernel: str):
    """HBM bytes per launch measured with rocprofv3 counters (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy kernel), from profiles/traffic.json."""
    p = ROOT / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        ent = d.get("kernels", {}).get(kernel)
        if ent and d.get("workload") == "udf_201x201x200":
            return float(ent["hbm_bytes_per_launch"])
    except (ValueError, KeyError, TypeError):
        return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the CPU baseline sample (default 150, ~10-20 s of CPU work)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost ratio")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 32-run ensemble")
    ap.add_argument("--cpu-steps-1core", type=int, default=12,
                    help="steps of the one-thread CPU baseline sample")
    ap.add_argument("--no-hernquist", action="store_true",
                    help="skip the 513x513x256 Hernquist forward and adjoint leg (~1.5 min)")
    ap.add_argument("--hq-steps", type=int, default=HQ_STEPS,
                    help="steps of the Hernquist adjoint leg (shorter runs for profiling)")
    ap.add_argument("--hq-ck", type=int, default=HQ_CK, help="its checkpoint interval")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batch of 8 Hernquist adjoint runs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import vlasolve_amd as V

    solver = V.Solver(WORKLOAD)
    ctx = solver.context
    npts = solver.nR * solver.nU * solver.nJ
    solver.step(args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    solver.step(args.steps)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    value = npts * args.steps / wall
    # Per-kernel durations: for each kernel of the step, the same number of
    # steps again with a HIP event on the solver's stream just before and
    # just after each of its launches (the other kernels run as in the timed
    # steps above).
    per = {}
    for name in STEP_KERNELS:
        ctx.set_timing(True, kernel=name)
        ctx.synchronize()
        solver.step(args.steps)
        ctx.synchronize()
        times = ctx.kernel_times(cap=2 * (args.steps + 1))
        if times:
            per[name] = (sum(ms for _, ms in times), len(times))
    ctx.set_timing(False)
    if args.calibrate_copy:
        V.copy_probe(args.calibrate_copy << 20)

    dominant = max((k for k in per if algorithmic_bytes(k, 1, 1, 1)), key=lambda k: per[k][0])
    tot, cnt = per[dominant]
    avg_s = tot / cnt / 1e3
    alg = algorithmic_bytes(dominant, solver.nR, solver.nU, solver.nJ)
    achieved = alg / avg_s / 1e9
    step_alg = STEP_BYTES_PER_POINT * npts
    step_achieved = step_alg / (wall / args.steps) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic_for(dominant), "kernel": dominant,
                "avg_launch_us": round(avg_s * 1e6, 2), "algorithmic_bytes_per_launch": alg,
                "kernels_us_per_step": {k: round(v[0] / v[1] * 1e3, 2)
                                        for k, v in sorted(per.items())},
                "step": {"algorithmic_bytes": step_alg, "achieved": round(step_achieved, 1),
                         "frac": round(step_achieved / HBM_PEAK_GBS, 4),
                         "basis": "72 B per point per step (moment 8, kick 16, r-solve 16, "
                                  "v-solve 16, drift 16) over ms_per_step"}}

    adj = None
    if not args.no_adjoint:
        try:
            adj = adjoint_ratio(V)
        except Exception as e:  # reported, not fatal for the forward figure
            adj = {"ratio": None, "error": str(e)[:200]}

    ens = None
    if not args.no_ensemble:
        try:
            ens = ensemble_speedup(V)
        except Exception as e:  # reported, not fatal for the forward figure
            ens = {"speedup": None, "error": str(e)[:200]}

    hq = None
    if not args.no_hernquist:
        try:
            hq = adjoint_ratio_hernquist(V, args.hq_steps, args.hq_ck)
        except Exception as e:  # reported, not fatal for the forward figure
            hq = {"ratio": None, "error": str(e)[:200]}

    hqb = None
    if not args.no_hernquist and not args.no_batch:
        try:
            hqb = adjoint_batch_ratio(V)
        except Exception as e:  # reported, not fatal for the forward figure
            hqb = {"ratio": None, "error": str(e)[:200]}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_steps or 150, args.cpu_steps_1core)
        except Exception as e:  # reported, not fatal for the GPU figure
            cpu = {"value": None, "error": str(e)[:200]}

    line = {
        "metric": "phase-space cell-updates/s at 200^3 FP64 (vs CPU); adjoint/forward cost ratio",
        "value": value, "unit": "point-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "UDF sphere, 201 log-r x 201 v x 200 quadratic-j (Pres 200, "
                               "Ures 200 200), dt 0.005, FP64, delayed central kernel",
                   "points": npts, "global_batch": 1, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "adjoint": adj,
        "ensemble": ens,
    }
    if hq is not None:
        line["hernquist"] = hq
    if hqb is not None:
        line["hernquist_batch"] = hqb
    print(json.dumps(line))


if __name__ == "__main__":
    main()
