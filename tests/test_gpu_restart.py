"""Restart through the reference's file format on the GPU: a run that saves
its state with SAVE_FLOW's layout at a step boundary and a new solver that
reads it with READ_FLOW's (ekman_bl_amd/restart.py) continue identically.
Config 1 with a fixed time step: the restart files do not carry DELTA_T
(READ_FLOW takes it from input.dat, src/diablo.F90:918-1000), and the plain
projection starts every solve from zero, so no other state is carried."""
import numpy as np
import pytest

from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import config

pytestmark = pytest.mark.gpu


def test_restart_files_continue_bit_identically(tmp_path):
    cfg = config(1)
    p = cfg.params
    p.variable_dt = 0
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    a = S.DuctSolver(p, xp, yp)
    a.set_initial_physical(*ic)
    a.time_step(2)
    paths = a.save_flow(str(tmp_path))
    print("[restart] saved", paths, flush=True)
    a.time_step(2)
    sa = a.get_state()
    da = a.diag()
    a.close()
    b = S.DuctSolver(p, xp, yp)
    fl = b.read_flow(str(tmp_path))
    assert fl.time_step == 2 and fl.time > 0.0
    b.time_step(2)
    sb = b.get_state()
    db = b.diag()
    b.close()
    print(f"[restart] time {da.time} vs {db.time}, step {da.time_step} vs {db.time_step}", flush=True)
    assert da.time_step == db.time_step == 4 and da.time == db.time
    for x, y in zip(sa[:4], sb[:4]):
        assert np.array_equal(x, y)
