#!/usr/bin/env python3
"""Benchmark of the duct time step on one MI355X.

Workload: the stratified Ekman layer NX=256, NY+2=193 (tanh stretched),
NZ+2=257, Re_delta=1000, Pr=0.7, N/f=30, seeded Ekman spiral plus 5% noise,
theta with 1% noise.  One step = three RKW3+CN substeps, each with the full
per-kx pressure projection.  Metric: grid-point-steps per second
(Mpt-steps/s, N_pts = NX*(NY+2)*(NZ+2) = 12,697,856) at ||div u|| < 1e-10.

The headline form (`value`) is the consistent projection with FP32
alternating-line V(1,1) cycles inside FP64 residual iterations
(proj_mode 2, max|residual| <= 1e-12 per real solve, each solve started
from zero: a V(1,1) cycle reduces the residual about a thousandfold, so
the previous substep's correction as a start saves no cycle); `forms` also times
the same projection with FP64 cycles (proj_mode 1) and the plain MUDPACK
replica (proj_mode 0, the reference's W(2,1) cycles to relmax 1e-6, which
does not reach the divergence bound).  The divergence columns are the
reference's diagnostic (mean |J div u| over interior points, and its max).

`roofline` is for the dominant kernel of the headline form (the batched
projection solve, one launch per substep), timed live with HIP events on the
solver's stream; algorithmic bytes from ekman_bl_amd/roofline.py.
`cpu_baseline` is the CPU oracle (the reference algorithm, proj_mode 0) on
the host's cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ekman_bl_amd import roofline  # noqa: E402
from ekman_bl_amd import solver as S  # noqa: E402
from ekman_bl_amd.grid import tanh_grid  # noqa: E402
from ekman_bl_amd.ic import ekman_initial  # noqa: E402
from ekman_bl_amd.params import config  # noqa: E402

HBM_PEAK_GBS = 8000.0
DIV_BOUND = 1e-10
FORM_NAMES = {0: "plain FP64 MUDPACK-replica W(2,1), relmax 1e-6",
              1: "consistent projection, FP64 alternating-line V(1,1)",
              2: "consistent projection, FP32 alternating-line V(1,1) cycles in FP64 outer iterations"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_form(cfg, xp, yp, ic, mode: int, steps: int, warmup: int, tol: float, guess: int = 0, krylov: int = 1):
    p = cfg.params.copy()
    p.proj_mode = mode
    p.proj_tol = tol
    p.proj_maxit = 50
    p.proj_guess = guess if mode > 0 else 0
    p.proj_krylov = krylov if mode > 0 else 0
    s = S.DuctSolver(p, xp, yp)
    s.set_initial_physical(*ic)
    for _ in range(warmup):
        s.time_step(1)
    s.sync()
    S.prof_enable(s, True)
    t0 = time.perf_counter()
    cyc = 0
    div_ref, div_max = [], []
    for _ in range(steps):
        for _q in range(3):
            s.rk_duct()
            d = s.diag()
            cyc += d.mg_cycles
            div_ref.append(d.div_ref)
            div_max.append(d.div_max)
    s.sync()
    el = time.perf_counter() - t0
    mg_ms, launches, solves, cycles = S.prof_get(s)
    dt = s.diag().delta_t
    s.close()
    return {"mode": mode, "el": el, "steps": steps, "cyc": cyc, "div_ref": max(div_ref), "div_max": max(div_max),
            "mg_ms": mg_ms, "launches": launches, "solves": solves, "cycles": cycles, "dt": dt}


def workload_name(cfg, mode: int, tol: float, krylov: int) -> str:
    return (f"{cfg.name}: RK3+CN step"
            + (", dynamic Smagorinsky LES of velocity and scalar" if cfg.params.les_model else "")
            + f", {FORM_NAMES[mode]}"
            + (f", max|residual| <= {tol:g}" if mode else "")
            + (", GCR(1) outer iteration" if mode and krylov else ""))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_tight_solve_hbm_traffic.md")


def measured_traffic(workload: str):
    """HBM bytes per launch of the headline solve kernel, as measured with
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (scaled by the copy-kernel
    calibration) and recorded in the traffic summary: (bytes, cycles per
    launch of the measured dispatches), or None when the summary does not
    cover this workload."""
    try:
        kv = {}
        for line in open(TRAFFIC_FILE):
            if ":" in line and line.startswith("- `"):
                k, v = line[3:].split("`", 1)
                kv[k] = v.split(":", 1)[1].strip().strip("`")
        if kv.get("workload") != workload:
            return None
        return int(kv["traffic_bytes_per_launch"]), float(kv["cycles_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def cpu_baseline(cfg, xp, yp, ic, budget_s: float):
    """CPU oracle (clean-room FP64 restatement of the reference algorithm,
    OpenMP over the host cores) on the same workload: whole steps after the
    initial projection until the time budget is spent (at least one)."""
    from oracle import oracle as O
    O.lib()
    p = cfg.params.copy()
    p.proj_mode = 0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(p, xp, yp)
    o.set_initial_physical(*ic)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.time_step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or el / steps * (steps + 1) > 2.0 * budget_s:
            break
    el = time.perf_counter() - t0
    value = p.npts * steps / el / 1e6
    return {"value": round(value, 4), "unit": "Mpt-steps/s", "cores": cores, "kind": "port",
            "sample": f"{cfg.name}, reference algorithm (MUDPACK W(2,1) projection): {steps} step(s) "
                      f"after the initial projection, {el:.1f} s wall"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--mode", type=int, default=2, help="headline projection form (0, 1, 2)")
    ap.add_argument("--tol", type=float, default=1e-12, help="max |residual| per real solve, modes 1 and 2")
    ap.add_argument("--guess", type=int, default=0, choices=(0, 1),
                    help="warm start of modes 1 and 2 (0: from zero, 1: previous substep's correction)")
    ap.add_argument("--krylov", type=int, default=1, choices=(0, 1),
                    help="outer iteration of modes 1 and 2 (0: plain refinement, 1: GCR(1))")
    ap.add_argument("--side-steps", type=int, default=5, help="timed steps of the other forms (0: skip)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-oracle work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="run the plain copy kernel on this many MiB once (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the duct solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(the duct is not sharded across GPUs)", file=sys.stderr)
        return 2
    cfg = config(args.config)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    L = S.lib()
    if args.calib_copy:
        n = args.calib_copy * (1 << 20) // 8
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        L.ekm_dev_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.ekm_copy_dev.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.ekm_dev_free.argtypes = [ctypes.c_void_p]
        L.ekm_dev_alloc(ctypes.byref(a), n * 8)
        L.ekm_dev_alloc(ctypes.byref(b), n * 8)
        L.ekm_copy_dev(b, a, n, None)
        L.ekm_dev_free(a)
        L.ekm_dev_free(b)
        log(f"copy kernel: {n * 8} bytes read, {n * 8} bytes written")
    log(f"[bench] {cfg.name}: headline form {args.mode} ({FORM_NAMES[args.mode]}), {args.steps} steps")
    h = run_form(cfg, xp, yp, ic, args.mode, args.steps, args.warmup, args.tol, args.guess, args.krylov)
    value = p.npts * args.steps / h["el"] / 1e6
    ms_per_step = h["el"] / args.steps * 1e3
    nsolve = 2 * (p.nx // 3 + 1)
    avg_ms = h["mg_ms"] / max(h["launches"], 1)
    cyc_per_launch = h["cycles"] / max(h["launches"], 1)
    if args.mode == 0:
        cbar = h["cycles"] / max(h["solves"], 1)
        launch_bytes = roofline.mg_launch_bytes(p, nsolve, cbar)
        step_bytes = roofline.step_bytes(p, cbar)
        kname = "k_mg_solve (batched MUDPACK replica, one launch per projection)"
    else:
        width = 4 if args.mode == 2 else 8
        launch_bytes = roofline.tight_launch_bytes(p, nsolve, cyc_per_launch, width, args.krylov)
        step_bytes = roofline.tight_step_bytes(p, cyc_per_launch, width, args.krylov)
        kname = "k_tight_solve (batched consistent projection, one launch per projection)"
    achieved = launch_bytes / (avg_ms * 1e-3) / 1e9
    log(f"[bench] {ms_per_step:.2f} ms/step, {value:.1f} Mpt-steps/s, solve {avg_ms:.3f} ms/launch, "
        f"cycles/step {h['cyc'] / args.steps:.0f}, div_ref max {h['div_ref']:.3e}, div_max max {h['div_max']:.3e}")
    forms = {}
    for mode in (0, 1, 2):
        if mode == args.mode:
            r = h
        elif args.side_steps > 0:
            log(f"[bench] form {mode} ({FORM_NAMES[mode]}), {args.side_steps} steps")
            r = run_form(cfg, xp, yp, ic, mode, args.side_steps, 1, args.tol, args.guess, args.krylov)
        else:
            continue
        forms[str(mode)] = {
            "form": FORM_NAMES[mode],
            "mpt_steps_per_s": round(p.npts * r["steps"] / r["el"] / 1e6, 3),
            "ms_per_step": round(r["el"] / r["steps"] * 1e3, 3),
            "mg_cycles_per_step": round(r["cyc"] / r["steps"], 2),
            "cycle_type": "W(2,1) point GS" if mode == 0 else "V(1,1) alternating zebra line GS",
            "solve_ms_per_launch": round(r["mg_ms"] / max(r["launches"], 1), 4),
            "div_ref_max": r["div_ref"],
            "div_max_max": r["div_max"],
            "meets_div_bound": bool(r["div_max"] < DIV_BOUND),
        }
    workload = workload_name(cfg, args.mode, args.tol, args.krylov)
    tr = measured_traffic(workload) if args.mode == 2 else None
    out = {
        "metric": "Mpt-steps/s at ||div u||<1e-10",
        "value": round(value, 3),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.mode != 2 else "f64 (projection cycles f32)",
        "data": "synthetic: seeded laminar Ekman spiral + 5% uniform noise (splitmix64), theta 1% noise",
        "config": {"workload": workload,
                   "nx": p.nx, "n_y": p.ny + 2, "n_z": p.nz + 2, "n_pts": p.npts},
        "div_ref_max": h["div_ref"],
        "div_max_max": h["div_max"],
        "meets_div_bound": bool(h["div_max"] < DIV_BOUND),
        "mg_cycles_per_step": round(h["cyc"] / args.steps, 2),
        "dt_final": h["dt"],
        "step_roofline_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": tr[0] if tr else None,
                     "traffic_cycles_per_launch": tr[1] if tr else None,
                     "traffic_source": "profiles/k_tight_solve_hbm_traffic.md" if tr else None,
                     "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(launch_bytes)},
        "forms": forms,
    }
    if not args.no_cpu:
        log("[bench] CPU oracle baseline")
        out["cpu_baseline"] = cpu_baseline(cfg, xp, yp, ic, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
