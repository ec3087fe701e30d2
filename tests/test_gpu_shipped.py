"""The reference run's own configuration on the GPU: rundir_unst's input.dat
and grid_def (fixtures: tests/golden/input_unst.dat, grid_yz_unst.json;
256 x 127 x 319, the GRID_YZ.dat tanh family with gamma 2.3) with the type-6
z rows, the scalar's type-5 bottom, the vertical sponge (NY_S_T = 120), the
dynamic mean-flow forcing (F_TYPE 1) and MUDPACK_BC 0 0 2 2: periodic
multigrid z ends and no specified side, so the kx = 0 pressure problem is
singular.  The initial field is the seeded Ekman spiral at the run's
velocity scale (WBULK0 = 0.05) with 5% noise (the run's DIABLO.START is not
shipped), and the background scalar a stable linear profile (the run reads
TH_BAR from its restart).

The consistent projection (FP32 cycles, GCR(1)) must hold max |J div u|
below 1e-10 after the initial projection and every substep of two steps,
with every solve converged, the singular kx = 0 solve included (its
right-hand side made compatible with the operator's left null vector); the
plain MUDPACK replica, the reference's algorithm, runs beside it and the
distance between the two forms' fields is reported."""
import os

import numpy as np
import pytest

from ekman_bl_amd import solver as S
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import params_from_fixture

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def shipped_params():
    return params_from_fixture(os.path.join(GOLDEN, "grid_yz_unst.json"), os.path.join(GOLDEN, "input_unst.dat"))


def test_shipped_configuration_consistent_projection():
    p, xp, yp = shipped_params()
    assert list(p.mudpack_bc) == [0, 0, 2, 2]
    ic = ekman_initial(p, yp, 1, 0.05, 0.01, ug=p.wbulk0)
    tb = 1e-3 * yp
    ph = p.copy()
    ph.proj_mode, ph.proj_tol, ph.proj_maxit, ph.proj_krylov = 2, 1e-12, 50, 1
    s = S.DuctSolver(ph, xp, yp, th_bar=tb)
    s0 = S.DuctSolver(p, xp, yp, th_bar=tb)
    for x in (s, s0):
        x.set_initial_physical(*ic)
    d = s.diag()
    cyc = s.solve_cycles()
    print(f"[shipped] initial projection: div_max {d.div_max:.2e}, cycles kx=0 {cyc[:2].tolist()}, "
          f"others max {cyc[2:].max()}", flush=True)
    assert d.status == 0 and d.div_max < 1e-10 and d.mg_unconverged == 0
    for sub in range(6):
        s.rk_duct()
        s0.rk_duct()
        d, d0 = s.diag(), s0.diag()
        cyc = s.solve_cycles()
        print(f"[shipped] substep {sub}: div_max {d.div_max:.2e} (replica {d0.div_max:.2e}), cycles {d.mg_cycles} "
              f"(kx=0 {cyc[:2].tolist()}), replica W cycles {d0.mg_cycles}, dt {d.delta_t:.4f}", flush=True)
        assert d.status == 0 and d0.status == 0
        assert d.div_max < 1e-10 and d.mg_unconverged == 0
    rel = [float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(s.get_state(), s0.get_state())]
    print("[shipped] consistent vs replica fields (rel L2):", " ".join(f"{v:.2e}" for v in rel), flush=True)
    assert all(np.isfinite(rel))
    for x in (s, s0):
        x.close()
