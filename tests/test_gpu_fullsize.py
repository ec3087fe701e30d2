"""The headline form at the full sizes of the 512x257x513 and 1024x385x769
stratified cases, which the oracle cannot run: properties that do not depend
on the size.  Every substep of one step reports status 0, max |J div u|
below 1e-10 and a positive cycle count; the x statistics (means, rms sums,
fluxes) stay finite; the mean of U1 at the top rows stays near the
geostrophic value the initial condition sets.  The 1024x385x769 case then
restarts from that state at step 3 (a restart's first step refreshes the
subgrid coefficients in every substep) and runs a step with both subgrid
models active: status 0, max |J div u| below 1e-10, C_DYN >= 0,
C_DYN_TH >= 0, NU_T and KAPPA_T finite and non-negative, some NU_T
positive."""
import numpy as np
import pytest

from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import YMAX, config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 5])
def test_full_size_headline_form_properties(n):
    cfg = config(n)
    p = cfg.params.copy()
    p.proj_mode, p.proj_tol, p.proj_maxit, p.proj_krylov, p.proj_guess = 2, 1e-12, 50, 1, 0
    print(f"[fullsize] config {n}: {p.nx}x{p.ny + 2}x{p.nz + 2}, building the initial condition", flush=True)
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    s = S.DuctSolver(p, xp, yp)
    s.set_initial_physical(*ic)
    del ic
    d = s.diag()
    print(f"[fullsize] config {n}: initial projection div_max {d.div_max:.2e}", flush=True)
    assert d.div_max < 1e-10
    for sub in range(3):
        s.rk_duct()
        d = s.diag()
        print(f"[fullsize] config {n} substep {sub}: status {d.status} div_max {d.div_max:.2e} "
              f"cycles {d.mg_cycles} (max {d.mg_max_cycles})", flush=True)
        assert d.status == 0
        assert d.div_max < 1e-10
        assert d.mg_cycles > 0 and d.mg_max_cycles < p.proj_maxit
    st = s.x_stats()
    for name, plane in st.items():
        assert np.all(np.isfinite(plane)), name
    ug = p.u_bc_c1[0][YMAX]
    top = st[S.DuctSolver.XSTATS[0]][p.ny - 2:p.ny, 2:p.nz]
    assert np.max(np.abs(top - ug)) < 0.1 * max(abs(ug), 1.0)
    if p.les_model:
        state = s.get_state()
        s.set_state(*state[:4], state[4], time=0.0, time_step=3)
        del state
        for sub in range(3):
            s.rk_duct()
            d = s.diag()
            print(f"[fullsize] config {n} LES substep {sub}: status {d.status} div_max {d.div_max:.2e} "
                  f"cycles {d.mg_cycles}", flush=True)
            assert d.status == 0 and d.div_max < 1e-10
        nut, cd = s.get_les()
        kap, cdt = s.get_les_th()
        print(f"[fullsize] config {n} LES: max NU_T {nut.max():.3e}, max C_DYN {cd.max():.3e}, "
              f"max KAPPA_T {kap.max():.3e}, max C_DYN_TH {cdt.max():.3e}", flush=True)
        assert np.all(np.isfinite(nut)) and np.all(np.isfinite(kap))
        assert cd.min() >= 0.0 and cdt.min() >= 0.0 and nut.min() >= 0.0 and kap.min() >= 0.0
        assert nut.max() > 0.0
        del nut, kap
    s.close()
