#!/usr/bin/env python3
"""Benchmark of the plain FP64 duct time step on one MI355X.

Workload: the stratified Ekman-layer configuration NX=256, NY+2=193 (tanh
stretched), NZ+2=257, Re_delta=1000, Pr=0.7, N/f=30, seeded Ekman spiral plus
5% noise; one step = three RKW3+CN substeps, each with the full per-kx
multigrid projection.  Metric: grid-point-steps per second (Mpt-steps/s),
N_pts = NX*(NY+2)*(NZ+2) = 12,697,856.

Prints one JSON line with the contract fields plus `roofline` for the
dominant kernel (the batched multigrid launch, timed live with HIP events on
its stream; algorithmic bytes from the byte contract in
ekman_bl_amd/roofline.py) and `cpu_baseline` (the CPU oracle on the host's
cores over a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ekman_bl_amd import roofline  # noqa: E402
from ekman_bl_amd import solver as S  # noqa: E402
from ekman_bl_amd.grid import tanh_grid  # noqa: E402
from ekman_bl_amd.ic import ekman_initial  # noqa: E402
from ekman_bl_amd.params import config  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cfg, xp, yp, u1, u2, u3, th, budget_s: float):
    """CPU oracle (clean-room FP64 restatement, OpenMP over the host cores) on
    the same workload: the initial projection plus whole steps until the
    time budget is spent (at least one)."""
    from oracle import oracle as O
    O.lib()
    p = cfg.params
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = O.Oracle(p, xp, yp)
    o.set_initial_physical(u1, u2, u3, th)
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
            "sample": f"{cfg.name}: {steps} step(s) after the initial projection, {el:.1f} s wall"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-oracle work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="run the plain copy kernel on this many MiB once (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the duct is not sharded across GPUs)", file=sys.stderr)
        return 2
    cfg = config(args.config)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    u1, u2, u3, th = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
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
    log(f"[bench] {cfg.name}: creating solver")
    s = S.DuctSolver(p, xp, yp)
    s.set_initial_physical(u1, u2, u3, th)
    log(f"[bench] warmup {args.warmup} steps")
    for _ in range(args.warmup):
        s.time_step(1)
    s.sync()
    S.prof_enable(s, True)
    log(f"[bench] timing {args.steps} steps")
    t0 = time.perf_counter()
    cyc = 0
    div_ref, div_max = [], []
    for _ in range(args.steps):
        for _q in range(3):
            s.rk_duct()
            d = s.diag()
            cyc += d.mg_cycles
            div_ref.append(d.div_ref)
            div_max.append(d.div_max)
    s.sync()
    el = time.perf_counter() - t0
    mg_ms, mg_launches, mg_solves, mg_cycles = S.prof_get(s)
    dgs = s.diag()
    value = p.npts * args.steps / el / 1e6
    ms_per_step = el / args.steps * 1e3
    cbar = mg_cycles / max(mg_solves, 1)
    mg_bytes = roofline.mg_launch_bytes(p, mg_solves / max(mg_launches, 1), cbar)
    avg_ms = mg_ms / max(mg_launches, 1)
    achieved = mg_bytes / (avg_ms * 1e-3) / 1e9
    step_bytes = roofline.step_bytes(p, cbar)
    log(f"[bench] {ms_per_step:.2f} ms/step, {value:.1f} Mpt-steps/s, mg {avg_ms:.3f} ms/launch, "
        f"cbar {cbar:.2f}, div_ref max {max(div_ref):.3e}, div_max max {max(div_max):.3e}")
    out = {
        "metric": "Mpt-steps/s (plain FP64 step)",
        "value": round(value, 3),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded laminar Ekman spiral + 5% uniform noise (splitmix64), theta 1% noise",
        "config": {"workload": cfg.name + " plain FP64 RK3+CN step, MUDPACK-replica W(2,1) multigrid "
                   f"tolmax {p.mg_tolmax:g} maxcy {p.mg_maxcy}",
                   "nx": p.nx, "n_y": p.ny + 2, "n_z": p.nz + 2, "n_pts": p.npts},
        "mg_cycles_per_step": round(cyc / args.steps, 2),
        "mg_cycles_per_solve": round(cbar, 3),
        "mg_cycle_type": "W(2,1)",
        "div_ref_max": max(div_ref),
        "div_max_max": max(div_max),
        "dt_final": dgs.delta_t,
        "step_roofline_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": {"kernel": "k_mg_solve (batched multigrid, one launch per projection)",
                     "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(mg_bytes)},
    }
    s.close()
    if not args.no_cpu:
        log("[bench] CPU oracle baseline")
        out["cpu_baseline"] = cpu_baseline(cfg, xp, yp, u1, u2, u3, th, args.cpu_budget)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
