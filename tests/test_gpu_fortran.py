"""The Fortran host (ekman_bl_amd/fortran/duct_driver, ISO_C_BINDING) drives
the GPU step the way the reference's main program drives RK_DUCT
(src/diablo.F90:58-190).  Its final state must equal the state of the same
library driven through ctypes bit for bit, and the CPU oracle's to 1e-10.
"""
import os
import subprocess

import numpy as np
import pytest

from ekman_bl_amd import build as B
from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid, write_grid_yz
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import config

pytestmark = pytest.mark.gpu

NSTEPS = 2


def _namelist(p, n_steps):
    lines = ["&duct"]
    for name, _ in p._fields_:
        v = getattr(p, name)
        if name == "reserved":
            continue
        if name in ("u_bc", "u_bc_c1"):
            # C [3][4] is Fortran (4,3): element (slot+1, comp+1)
            for comp in range(3):
                for slot in range(4):
                    lines.append(f"  p%{name}({slot + 1},{comp + 1}) = {v[comp][slot]!r}")
        elif hasattr(v, "__len__"):
            for i in range(len(v)):
                lines.append(f"  p%{name}({i + 1}) = {v[i]!r}")
        else:
            lines.append(f"  p%{name} = {v!r}")
    lines.append(f"  n_steps = {n_steps}")
    lines.append("/")
    return "\n".join(lines).replace("e-", "d-").replace("e+", "d+") + "\n"


def test_fortran_driver_matches_ctypes_and_oracle(tmp_path, oracle_lib):
    if not os.path.exists(B.DRIVER):
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    u1, u2, u3, th = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    run = tmp_path / "run"
    run.mkdir()
    (run / "duct.nml").write_text(_namelist(p, NSTEPS))
    write_grid_yz(str(run / "GRID_YZ.dat"), xp, yp)
    with open(run / "ic.bin", "wb") as f:
        for a in (u1, u2, u3, th):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    print("[fortran] running duct_driver", flush=True)
    r = subprocess.run([B.DRIVER, str(run)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    raw = np.fromfile(run / "state.bin", dtype=np.complex128)
    nf = int(np.prod(p.shape_fourier))
    assert raw.size == 5 * nf
    fstate = [raw[i * nf:(i + 1) * nf].reshape(p.shape_fourier) for i in range(5)]

    # the same library through ctypes, from the grid file the driver read
    gx, gy = np.loadtxt(run / "GRID_YZ.dat").T
    s = S.DuctSolver(p, gx.reshape(xp.shape), gy.reshape(yp.shape))
    s.set_initial_physical(u1, u2, u3, th)
    s.time_step(NSTEPS)
    cstate = s.get_state()
    s.close()
    for a, b in zip(fstate, cstate):
        assert np.array_equal(a, b)

    O = oracle_lib
    o = O.Oracle(p, gx.reshape(xp.shape), gy.reshape(yp.shape))
    o.set_initial_physical(u1, u2, u3, th)
    o.time_step(NSTEPS)
    for a, b in zip(fstate, o.get_state()):
        n = np.linalg.norm(b)
        assert np.linalg.norm(a - b) / (n if n > 0 else 1.0) < 1e-10
