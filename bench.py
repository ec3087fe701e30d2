#!/usr/bin/env python3
"""Benchmark of the duct time step on one MI355X.

Workload (default): the stratified Ekman layer with the dynamic Smagorinsky
subgrid models of the velocity and the scalar, NX=1024, NY+2=385 (tanh
stretched), NZ+2=769, Re_delta=5000, Pr=0.7, N/f=50, seeded Ekman spiral plus
5% noise, theta with 1% noise.  One step = three RKW3+CN substeps, each with
the full per-kx pressure projection.  Metric: grid-point-steps per second
(Mpt-steps/s, N_pts = NX*(NY+2)*(NZ+2) = 303,170,560) at ||div u|| < 1e-10,
with the multigrid cycles per step.  The warm-up covers step 1 of the new
flow, on which the reference does not run the subgrid models, so the models
run in every timed step.

Forms (each from the same seeded state, each checked against the bound):
- headline (`value`): the consistent projection with FP32 alternating-line
  V(1,1) cycles inside FP64 GCR(1) outer iterations (proj_mode 2, max|residual|
  <= 1e-12 per real solve), all solves of a projection and all ADI systems of
  a substep batched; launched eagerly so that the projection's solve kernel is
  timed with HIP events on the solver's stream (`roofline`);
- the same replayed from captured HIP graphs (`forms.graph`);
- the same started from the learned guess (`forms.learned`, when trained
  weights are found);
- FP64 cycles (proj_mode 1): the reference for the per-field errors;
- the plain FP64 MUDPACK replica (proj_mode 0, the reference's W(2,1) cycles
  to relmax 1e-6): the drop-in, which does not reach the bound.
Per-field errors: relative L2 of U1, U2, U3, P, TH of each form against the
FP64-cycle form after the same steps.
`cpu_baseline`: the CPU oracle (the reference algorithm, OpenMP over the
host's cores) on one substep of the same workload, restarted from the FP64
form's state so that the subgrid models run.
`reference_run`: the reference run's own configuration (rundir_unst,
256x127x319 with periodic multigrid z ends and no specified side), the plain
replica and the consistent projection side by side (not part of `value`).
"""
from __future__ import annotations

import argparse
import ctypes
import gc
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
from ekman_bl_amd.params import config, params_from_fixture  # noqa: E402

HBM_PEAK_GBS = 8000.0
DIV_BOUND = 1e-10
FORM_NAMES = {0: "plain FP64 MUDPACK-replica W(2,1), relmax 1e-6",
              1: "consistent projection, FP64 alternating-line V(1,1)",
              2: "consistent projection, FP32 alternating-line V(1,1) cycles in FP64 outer iterations"}
DATA = os.path.join(ROOT, "ekman_bl_amd", "data")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_params(cfg, mode: int, tol: float, krylov: int, guess: int = 0):
    p = cfg.params.copy()
    p.proj_mode = mode
    p.proj_tol = tol
    p.proj_maxit = 50
    p.proj_guess = guess if mode > 0 else 0
    p.proj_krylov = krylov if mode > 0 else 0
    return p


def tuning_of(cfg, mode: int):
    """The autotuner's cached kernel tuning of this workload and form on this
    device (ekman_bl_amd/data/tuning.json), or None (the defaults)."""
    from ekman_bl_amd import autotune as AT
    return AT.load(cfg.name, mode)


def run_form(cfg, xp, yp, ic, mode, steps, warmup, tol, krylov, graph=False, net=None, prof=False,
             keep_state=False):
    """warmup + steps steps of one form; the timed part is the `steps` steps."""
    p = make_params(cfg, mode, tol, krylov, 2 if net is not None else 0)
    s = S.DuctSolver(p, xp, yp, tuning=tuning_of(cfg, mode))
    if net is not None:
        net.attach(s)
    s.set_initial_physical(*ic)
    for _ in range(warmup):
        s.time_step(1, graph=graph)
    s.sync()
    if prof:
        S.prof_enable(s, True)
    cyc, div_ref, div_max = 0, [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        for _q in range(3):
            s.rk_duct(graph=graph)
            d = s.diag()
            cyc += d.mg_cycles
            div_ref.append(d.div_ref)
            div_max.append(d.div_max)
    s.sync()
    el = time.perf_counter() - t0
    out = {"mode": mode, "el": el, "steps": steps, "cyc": cyc, "div_ref": max(div_ref), "div_max": max(div_max),
           "dt": s.diag().delta_t, "status": s.diag().status}
    if prof:
        out["mg_ms"], out["launches"], out["solves"], out["cycles"] = S.prof_get(s)
    if keep_state:
        out["state"] = s.get_state()
    s.close()
    return out


def field_errors(state, ref):
    names = ("u1", "u2", "u3", "p", "th")
    return {n: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for n, a, b in zip(names, state, ref)}


def workload_name(cfg, mode: int, tol: float, krylov: int) -> str:
    return (f"{cfg.name}: RK3+CN step"
            + (", dynamic Smagorinsky LES of velocity and scalar" if cfg.params.les_model else "")
            + f", {FORM_NAMES[mode]}"
            + (f", max|residual| <= {tol:g}" if mode else "")
            + (", GCR(1) outer iteration" if mode and krylov else ""))


def traffic_file(cfg) -> str:
    return os.path.join(ROOT, "profiles", f"k_tight_solve_{cfg.name.split('_')[0]}_hbm_traffic.md")


def measured_traffic(cfg, workload: str):
    """HBM bytes per launch of the headline solve kernel, as measured with
    rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (scaled by the copy-kernel
    calibration) and recorded in the traffic summary of this configuration:
    (bytes, cycles per launch of the measured dispatches, file), or None when
    no summary covers this workload."""
    path = traffic_file(cfg)
    try:
        kv = {}
        for line in open(path):
            if ":" in line and line.startswith("- `"):
                k, v = line[3:].split("`", 1)
                kv[k] = v.split(":", 1)[1].strip().strip("`")
        if kv.get("workload") != workload:
            return None
        return int(kv["traffic_bytes_per_launch"]), float(kv["cycles_per_launch"]), os.path.relpath(path, ROOT)
    except (OSError, KeyError, ValueError):
        return None


def learned_net(p, path: str | None):
    """A network for this plane with trained weights, or None."""
    from ekman_bl_amd import warmstart as WS
    if path is None:
        for cand in (os.path.join(DATA, f"wsnet_{p.nz + 2}x{p.ny + 2}.npz"), os.path.join(DATA, "wsnet.npz")):
            if os.path.exists(cand):
                path = cand
                break
    if path is None or not os.path.exists(path):
        return None, None
    if (p.nz + 1) % 32 or (p.ny + 1) % 32:
        return None, None
    net = WS.WarmStartNet(p.nz + 2, p.ny + 2, 2 * (p.nx // 3 + 1), trainable=False)
    net.load(path)
    return net, os.path.relpath(path, ROOT)


def reference_run(steps: int, tol: float, krylov: int) -> dict:
    """The reference run's own configuration (rundir_unst: input.dat and
    grid_def, kept as fixtures in tests/golden; 256x127x319, periodic
    multigrid z ends with no specified side): the plain replica and the
    consistent projection (FP32 cycles), one warm-up step and `steps` timed
    steps each from the seeded Ekman spiral at the run's velocity scale."""
    g = os.path.join(ROOT, "tests", "golden")
    p, xp, yp = params_from_fixture(os.path.join(g, "grid_yz_unst.json"), os.path.join(g, "input_unst.dat"))
    ic = ekman_initial(p, yp, 1, 0.05, 0.01, ug=p.wbulk0)
    tb = 1e-3 * yp  # background scalar (the run reads TH_BAR from its restart)
    out = {"config": f"rundir_unst {p.nx}x{p.ny + 2}x{p.nz + 2}, MUDPACK_BC {list(p.mudpack_bc)}, "
                     f"{p.npts} points; seeded Ekman spiral at WBULK0 scale, linear background scalar"}
    for mode, key in ((0, "plain_fp64_mudpack"), (2, "consistent_fp32_cycles")):
        q = p.copy()
        q.proj_mode = mode
        if mode:
            q.proj_tol, q.proj_maxit, q.proj_krylov = tol, 50, krylov
        s = S.DuctSolver(q, xp, yp, th_bar=tb)
        s.set_initial_physical(*ic)
        s.time_step(1)
        s.sync()
        cyc, dmax, unconv = 0, 0.0, 0
        t0 = time.perf_counter()
        for _ in range(steps):
            for _r in range(3):
                s.rk_duct()
                d = s.diag()
                cyc += d.mg_cycles
                dmax = max(dmax, d.div_max)
                unconv += d.mg_unconverged
        s.sync()
        el = time.perf_counter() - t0
        d = s.diag()
        out[key] = {"ms_per_step": round(el / steps * 1e3, 3), "mpt_steps_per_s": round(p.npts * steps / el / 1e6, 3),
                    "mg_cycles_per_step": round(cyc / steps, 2), "div_max_max": dmax,
                    "meets_div_bound": bool(dmax < DIV_BOUND), "unconverged_solves": unconv, "status": d.status}
        s.close()
    out["speedup"] = round(out["plain_fp64_mudpack"]["ms_per_step"] / out["consistent_fp32_cycles"]["ms_per_step"], 3)
    return out


def cpu_baseline(cfg, xp, yp, state):
    """CPU oracle (clean-room FP64 restatement of the reference algorithm,
    OpenMP over the host cores): one substep of the same workload, restarted
    from `state` so that the subgrid models run (a restart's first step
    recomputes their coefficients in every substep)."""
    from oracle import oracle as O
    O.lib()
    p = cfg.params.copy()
    p.proj_mode = 0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t_init = time.perf_counter()
    o = O.Oracle(p, xp, yp)
    o.set_state(*state[:4], state[4], time=0.0, time_step=2)
    t_init = time.perf_counter() - t_init
    t0 = time.perf_counter()
    o.rk_duct()
    el = time.perf_counter() - t0
    del o
    value = p.npts / 3.0 / el / 1e6
    return {"value": round(value, 4), "unit": "Mpt-steps/s", "cores": cores, "kind": "port",
            "sample": f"{cfg.name}, reference algorithm (MUDPACK W(2,1) projection, both subgrid models): one "
                      f"substep (a third of a step) restarted from the FP64 form's state, {el:.1f} s wall "
                      f"(set-up {t_init:.1f} s not counted)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=5)
    ap.add_argument("--mode", type=int, default=2, help="headline projection form (0, 1, 2)")
    ap.add_argument("--tol", type=float, default=1e-12, help="max |residual| per real solve, modes 1 and 2")
    ap.add_argument("--krylov", type=int, default=1, choices=(0, 1),
                    help="outer iteration of modes 1 and 2 (0: plain refinement, 1: GCR(1))")
    ap.add_argument("--side-steps", type=int, default=3, help="steps of the compared forms (0: skip them)")
    ap.add_argument("--side-warmup", type=int, default=4,
                    help="warm-up steps of the compared forms (4: the subgrid models' first refresh, at step 5, "
                         "falls in the timed steps)")
    ap.add_argument("--learned", default=None, help="trained warm-start weights (default: ekman_bl_amd/data)")
    ap.add_argument("--ref-run-steps", type=int, default=3,
                    help="timed steps of the reference run's own configuration, both forms (0: skip)")
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
    log(f"[bench] {cfg.name}: initial condition")
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
    # headline: eager launches, the solve kernel timed with HIP events
    log(f"[bench] headline form {args.mode} ({FORM_NAMES[args.mode]}): {args.warmup} warm-up + {args.steps} steps")
    h = run_form(cfg, xp, yp, ic, args.mode, args.steps, args.warmup, args.tol, args.krylov, prof=True)
    value = p.npts * args.steps / h["el"] / 1e6
    ms_per_step = h["el"] / args.steps * 1e3
    nsolve = 2 * (p.nx // 3 + 1)
    avg_ms = h["mg_ms"] / max(h["launches"], 1)
    cyc_per_launch = h["cycles"] / max(h["launches"], 1)
    if args.mode == 0:
        cbar = h["cycles"] / max(h["solves"], 1)
        launch_bytes = roofline.mg_launch_bytes(p, nsolve, cbar)
        kname = "k_mg_solve (batched MUDPACK replica, one launch per projection)"
    else:
        width = 4 if args.mode == 2 else 8
        launch_bytes = roofline.tight_launch_bytes(p, nsolve, cyc_per_launch, width, args.krylov)
        kname = "k_tight_solve (batched consistent projection, one launch per projection)"
    achieved = launch_bytes / (avg_ms * 1e-3) / 1e9
    log(f"[bench] {ms_per_step:.2f} ms/step, {value:.1f} Mpt-steps/s, solve {avg_ms:.3f} ms/launch, "
        f"cycles/step {h['cyc'] / args.steps:.0f}, div_max {h['div_max']:.3e}")
    gc.collect()
    forms = {}

    def add_form(key, r, desc, cycle_type, ref=None):
        forms[key] = {
            "form": desc,
            "mpt_steps_per_s": round(p.npts * r["steps"] / r["el"] / 1e6, 3),
            "ms_per_step": round(r["el"] / r["steps"] * 1e3, 3),
            "mg_cycles_per_step": round(r["cyc"] / r["steps"], 2),
            "cycle_type": cycle_type,
            "div_ref_max": r["div_ref"],
            "div_max_max": r["div_max"],
            "meets_div_bound": bool(r["div_max"] < DIV_BOUND),
            "status": r["status"],
        }
        if ref is not None and "state" in r:
            forms[key]["field_rel_l2_vs_fp64_form"] = field_errors(r["state"], ref)

    ref_state = None
    if args.side_steps > 0:
        sw, ss = args.side_warmup, args.side_steps
        vc = "V(1,1) alternating zebra line GS"
        log(f"[bench] form 1 ({FORM_NAMES[1]}): {sw} + {ss} steps")
        r1 = run_form(cfg, xp, yp, ic, 1, ss, sw, args.tol, args.krylov, keep_state=True)
        ref_state = r1.pop("state")
        add_form("fp64_cycles", r1, FORM_NAMES[1] + ", GCR(1)", vc)
        for key, kw, desc in (("fp32_cycles", {}, FORM_NAMES[2] + ", GCR(1)"),
                              ("graph", {"graph": True}, FORM_NAMES[2] + ", GCR(1), substeps replayed from HIP graphs")):
            log(f"[bench] {key}: {sw} + {ss} steps")
            r = run_form(cfg, xp, yp, ic, 2, ss, sw, args.tol, args.krylov, keep_state=True, **kw)
            add_form(key, r, desc, vc, ref_state)
            del r
            gc.collect()
        net, src = learned_net(p, args.learned)
        if net is not None:
            log(f"[bench] learned start ({src}): {sw} + {ss} steps")
            r = run_form(cfg, xp, yp, ic, 2, ss, sw, args.tol, args.krylov, net=net, keep_state=True)
            add_form("learned", r, FORM_NAMES[2] + f", GCR(1), solves started from the learned guess ({src})", vc,
                     ref_state)
            del r
            net.close()
            gc.collect()
        log(f"[bench] form 0 ({FORM_NAMES[0]}): {sw} + {ss} steps")
        r0 = run_form(cfg, xp, yp, ic, 0, ss, sw, args.tol, args.krylov, keep_state=True)
        add_form("plain_fp64_mudpack", r0, FORM_NAMES[0], "W(2,1) point GS", ref_state)
        del r0
        gc.collect()
    workload = workload_name(cfg, args.mode, args.tol, args.krylov)
    tr = measured_traffic(cfg, workload) if args.mode == 2 else None
    fp64 = forms.get("fp64_cycles")
    out = {
        "metric": "Mpt-steps/s at ||div u||<1e-10; MG V-cycles/step, learned+FP32 MG vs plain FP64",
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
        "config": {"workload": workload, "nx": p.nx, "n_y": p.ny + 2, "n_z": p.nz + 2, "n_pts": p.npts},
        "div_ref_max": h["div_ref"],
        "div_max_max": h["div_max"],
        "meets_div_bound": bool(h["div_max"] < DIV_BOUND),
        "mg_cycles_per_step": round(h["cyc"] / args.steps, 2),
        "cycle_type": "V(1,1) alternating zebra line GS" if args.mode else "W(2,1) point GS",
        "dt_final": h["dt"],
        "speedup_vs_fp64_cycles": round(fp64["ms_per_step"] / ms_per_step, 3) if fp64 else None,
        "speedup_vs_plain_fp64_mudpack": (round(forms["plain_fp64_mudpack"]["ms_per_step"] / ms_per_step, 3)
                                          if "plain_fp64_mudpack" in forms else None),
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": tr[0] if tr else None,
                     "traffic_cycles_per_launch": tr[1] if tr else None,
                     "traffic_source": tr[2] if tr else None,
                     "avg_launch_ms": round(avg_ms, 4), "cycles_per_launch": round(cyc_per_launch, 1),
                     "algorithmic_bytes_per_launch": int(launch_bytes)},
        "forms": forms,
        "tuning": tuning_of(cfg, args.mode),
    }
    if args.ref_run_steps > 0:
        log("[bench] the reference run's own configuration (rundir_unst), both forms")
        out["reference_run"] = reference_run(args.ref_run_steps, args.tol, args.krylov)
    if not args.no_cpu:
        log("[bench] CPU oracle baseline")
        st = ref_state
        if st is None:
            r = run_form(cfg, xp, yp, ic, 1, 1, 1, args.tol, args.krylov, keep_state=True)
            st = r["state"]
        out["cpu_baseline"] = cpu_baseline(cfg, xp, yp, st)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
