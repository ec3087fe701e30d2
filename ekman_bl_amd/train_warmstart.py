"""Train the learned initial guess of the pressure solves and measure what it
gains (the learned warm-start study on the 256x193x257 stratified grid).

    python -m ekman_bl_amd.train_warmstart --out DIR [--config 2] [--seeds 0-9]
        [--heldout 10-12] [--steps 4] [--epochs 0] [--train-seconds 150]

1. Pairs: for every training seed, one warm-up step of the headline form
   (consistent projection, FP32 cycles, GCR(1), starts from zero), then
   `--steps` steps whose projections are dumped: each real solve's RHS_TMP,
   the previous projection's solution of the same solve, kx, and the
   converged solution (2 (NX/3+1) solves per projection).
2. The input scale sigma (max|phi|/max|rhs|, median) from a first run.
3. Adam on the GPU with the library's forward, backward and update kernels,
   batch 32, cosine-decayed learning rate, for `--epochs` epochs or as many as
   fit in `--train-seconds`.
4. Held-out seeds: V-cycles per step and ms per step from zero, from the
   previous correction scaled by the RK weights (proj_guess 1), and from the
   network's guess (proj_guess 2); fields of each start against the zero
   start's (relative L2) and the divergence.
Writes DIR/wsnet.npz (parameters, sigma, kx_max) and DIR/warmstart.json.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import warmstart as WS
from .grid import tanh_grid
from .ic import ekman_initial
from .params import config


def _seeds(spec: str):
    a, _, b = spec.partition("-")
    return list(range(int(a), int(b or a) + 1))


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--seeds", default="0-9")
    ap.add_argument("--heldout", default="10-12")
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--eval-steps", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=0, help="0: as many as fit in --train-seconds")
    ap.add_argument("--train-seconds", type=float, default=150.0)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    cfg = config(args.config)
    p = cfg.params.copy()
    p.proj_mode, p.proj_krylov, p.proj_tol, p.proj_maxit = 2, 1, 1e-12, 50
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    nz2, ny2 = p.nz + 2, p.ny + 2
    nsolve = 2 * (p.nx // 3 + 1)
    kx_max = 2 * np.pi * (p.nx // 3) / p.lx
    train_seeds, held = _seeds(args.seeds), _seeds(args.heldout)
    per_run = 3 * args.steps * nsolve
    cap = per_run * len(train_seeds)
    net = WS.WarmStartNet(nz2, ny2, args.batch, trainable=True)
    log(f"[train] {cfg.name}: {nsolve} solves per projection, {cap} pairs from seeds {train_seeds}, "
        f"{net.nparams} parameters")

    def ic_of(seed):
        return ekman_initial(p, yp, seed, cfg.noise, cfg.th_noise)

    # input scale from a first run
    net.set_norm(1.0, kx_max)
    cal = WS.PairSet(net, per_run)
    WS.run_and_dump(p, xp, yp, ic_of(train_seeds[0]), cal, steps=1, warmup=1)
    sigma = WS.estimate_sigma(cal, cal.active(), net)
    del cal
    net.set_norm(sigma, kx_max)
    log(f"[train] sigma {sigma:.4e}, kx_max {kx_max:.4f}")
    pairs = WS.PairSet(net, cap)
    t0 = time.perf_counter()
    for sd in train_seeds:
        WS.run_and_dump(p, xp, yp, ic_of(sd), pairs, steps=args.steps, warmup=1)
        log(f"[train] seed {sd}: {pairs.n} pairs ({time.perf_counter() - t0:.1f} s)")
    idx = pairs.active()
    log(f"[train] {len(idx)} pairs with a nonzero right-hand side")
    # the first epoch is timed; the rest fit in --train-seconds (or --epochs)
    t1 = time.perf_counter()
    hist = WS.train(net, pairs, idx, epochs=1, batch=args.batch, lr=args.lr, log=log)
    t_ep = time.perf_counter() - t1
    epochs = args.epochs - 1 if args.epochs > 0 else max(0, int(args.train_seconds / max(t_ep, 1e-3)) - 1)
    if epochs > 0:
        hist += WS.train(net, pairs, idx, epochs=epochs, batch=args.batch, lr=args.lr, seed=1, log=log)
    train_s = time.perf_counter() - t1
    net.save(os.path.join(args.out, "wsnet.npz"))
    del pairs
    # held-out evaluation with an inference network sized for a projection
    inf = WS.WarmStartNet(nz2, ny2, nsolve, trainable=False)
    inf.set_params(net.get_params())
    inf.set_norm(sigma, kx_max)
    evals = []
    for sd in held:
        ic = ic_of(sd)
        r = {g: WS.eval_cycles(p, xp, yp, ic, inf if g == 2 else None, g, steps=args.eval_steps, warmup=1)
             for g in (0, 1, 2)}
        ref = r[0]["state"]
        row = {"seed": sd}
        for g, name in ((0, "zero"), (1, "scaled_previous"), (2, "learned")):
            rel = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(r[g]["state"], ref)]
            row[name] = {"cycles_per_step": r[g]["cycles_per_step"], "ms_per_step": round(r[g]["ms_per_step"], 3),
                         "div_max": r[g]["div_max"], "field_rel_vs_zero_start": max(rel)}
        log(f"[eval] seed {sd}: cycles/step zero {row['zero']['cycles_per_step']:.0f}, "
            f"scaled {row['scaled_previous']['cycles_per_step']:.0f}, learned {row['learned']['cycles_per_step']:.0f}; "
            f"ms/step {row['zero']['ms_per_step']:.2f} / {row['scaled_previous']['ms_per_step']:.2f} / "
            f"{row['learned']['ms_per_step']:.2f}")
        evals.append(row)
    mean = {k: float(np.mean([e[k]["cycles_per_step"] for e in evals])) for k in ("zero", "scaled_previous", "learned")}
    out = {"workload": cfg.name, "pairs": int(len(idx)), "train_seeds": train_seeds, "heldout_seeds": held,
           "params": net.nparams, "sigma": sigma, "epochs": len(hist), "train_seconds": round(train_s, 1),
           "epoch_seconds": t_ep, "loss_history": hist, "eval": evals, "mean_cycles_per_step": mean,
           "cycles_saved_per_step": mean["zero"] - mean["learned"]}
    with open(os.path.join(args.out, "warmstart.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("workload", "pairs", "epochs", "mean_cycles_per_step",
                                          "cycles_saved_per_step")}), flush=True)
    inf.close()
    net.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
