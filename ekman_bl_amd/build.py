"""Build the gfx950 library libekman_bl_amd.so in-tree with hipcc, and the
Fortran host (ISO_C_BINDING module + duct_driver) with amdflang when present.

Usage: python -m ekman_bl_amd.build [--jobs N]
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libekman_bl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "duct_driver")

SOURCES = ["ekm_fft.hip", "ekm_kernels.hip", "ekm_les.hip", "ekm_mg.hip", "ekm_mgt.hip", "ekm_wsnet.hip", "ekm_noalloc.hip", "ekm_solver.hip",
           "ekm_host.cpp"]

# -ffp-contract=off keeps every product and sum separately rounded, as the
# reference build and the CPU oracle do; packed FP32 is disabled (CDNA4
# packed-FP32 forms are not used by this double-precision code).
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 4, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


def build_fortran(verbose: bool = True) -> str:
    """ekman_bl_amd.mod + duct_driver, linked to the in-tree library with an
    $ORIGIN-relative run path."""
    mod_o = os.path.join(OBJ, "ekman_bl_amd_mod.o")
    drv_o = os.path.join(OBJ, "duct_driver.o")
    steps = [
        [FLANG, "-O2", "-J", OBJ, "-c", os.path.join(FORTRAN, "ekman_bl_amd_mod.f90"), "-o", mod_o],
        [FLANG, "-O2", "-I", OBJ, "-c", os.path.join(FORTRAN, "duct_driver.f90"), "-o", drv_o],
        [FLANG, drv_o, mod_o, "-L", HERE, "-lekman_bl_amd", "-Wl,-rpath,$ORIGIN/..", "-o", DRIVER + ".tmp"],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


if __name__ == "__main__":
    jobs = 4
    if "--jobs" in sys.argv:
        jobs = int(sys.argv[sys.argv.index("--jobs") + 1])
    build(jobs)
