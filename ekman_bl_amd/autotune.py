"""The project's autotuner of the solver's kernel parameters (ekm_tuning).

Every parameter it chooses is a launch or memory-placement parameter -- rows
per batch and workgroup size of the ADI sweeps, the replica's workgroup
size, the LDS cap of the consistent projection's coarse levels -- so the
solver's results are bit-identical for every choice (checked by
tests/test_gpu_autotune.py); only the time per step changes.

Search: coordinate descent from the defaults over the parameters that the
form uses (the replica's workgroup only for proj_mode 0, the subgrid-model
sweeps only with les_model), each candidate timed over whole steps of the
workload itself after one warm-up step, repeated `passes` times.  Results
are cached in a JSON file keyed by the workload and the device name.

    python -m ekman_bl_amd.autotune --config 2 [--mode 2] [--steps 2] [--cache PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

from . import solver as S
from .grid import tanh_grid
from .ic import ekman_initial
from .params import config

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "data", "tuning.json")

SPACE = {
    "adi_rows": (4, 8, 12),
    "adi_block": (64, 128, 256),
    "adi_les_rows": (4, 5, 6, 8),
    "mg_block": (256, 512, 1024),
    "tight_lds_kib": (0, 96, 48),
}


def device_name() -> str:
    """The current HIP device's name and architecture, from the library."""
    import ctypes
    L = S.lib()
    L.ekm_device_name.argtypes = [ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(256)
    if L.ekm_device_name(buf, 256) != 0:
        return "unknown"
    return buf.value.decode(errors="replace")


def cache_key(cfg_name: str, mode: int, device: str) -> str:
    return f"{cfg_name}|proj_mode {mode}|{device}"


def relevant(p) -> list:
    keys = ["adi_rows", "adi_block"]
    if p.les_model:
        keys.append("adi_les_rows")
    if p.proj_mode == 0:
        keys.append("mg_block")
    else:
        keys.append("tight_lds_kib")
    return keys


def time_candidate(p, xp, yp, ic, tuning: dict, steps: int) -> float:
    s = S.DuctSolver(p, xp, yp, tuning=tuning)
    s.set_initial_physical(*ic)
    s.time_step(1)          # warm-up, and the first step of a new flow runs no subgrid model
    s.sync()
    t0 = time.perf_counter()
    s.time_step(steps)
    el = time.perf_counter() - t0
    s.close()
    return el / steps * 1e3


def tune(p, xp, yp, ic, steps: int = 2, passes: int = 1, log=print) -> dict:
    best = S.default_tuning().as_dict()
    trials = []
    best_ms = min(time_candidate(p, xp, yp, ic, best, steps) for _ in range(passes))
    trials.append({"tuning": dict(best), "ms_per_step": round(best_ms, 3)})
    log(f"[autotune] default {best}: {best_ms:.2f} ms/step")
    for key in relevant(p):
        for val in SPACE[key]:
            if val == best[key]:
                continue
            cand = dict(best, **{key: val})
            ms = min(time_candidate(p, xp, yp, ic, cand, steps) for _ in range(passes))
            trials.append({"tuning": dict(cand), "ms_per_step": round(ms, 3)})
            log(f"[autotune] {key}={val}: {ms:.2f} ms/step")
            if ms < best_ms:
                best, best_ms = cand, ms
    return {"tuning": best, "ms_per_step": round(best_ms, 3), "trials": trials}


def load(cfg_name: str, mode: int, path: str = CACHE, device: str | None = None):
    """The cached tuning of this workload on this device, or None."""
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    ent = data.get(cache_key(cfg_name, mode, device or device_name()))
    return ent["tuning"] if ent else None


def save(cfg_name: str, mode: int, result: dict, path: str = CACHE, device: str | None = None) -> None:
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[cache_key(cfg_name, mode, device or device_name())] = result
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--mode", type=int, default=2)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--passes", type=int, default=1)
    ap.add_argument("--cache", default=CACHE)
    args = ap.parse_args()
    cfg = config(args.config)
    p = cfg.params.copy()
    p.proj_mode = args.mode
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    res = tune(p, xp, yp, ic, steps=args.steps, passes=args.passes,
               log=lambda m: print(m, file=sys.stderr, flush=True))
    save(cfg.name, args.mode, res, args.cache)
    print(json.dumps({"workload": cfg.name, "proj_mode": args.mode, **{k: res[k] for k in ("tuning", "ms_per_step")}}),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
