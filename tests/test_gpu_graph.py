"""The substep replayed from HIP graphs against the same substep launched
eagerly (ekm_rk_duct_graph_launch vs ekm_rk_duct), through the C ABI.

Both paths enqueue the same kernels with the same arguments (the substep's
scalar decisions -- DELTA_T from the CFL minimum, the mean-pressure
controller, the projection's checks -- run in one-thread device kernels), so
the states must agree bit for bit (the divergence diagnostic, an atomic sum,
to rounding).  The cases cover every variant the host
decides for a substep: the three RK stages, the CFL update, the subgrid
models with and without their coefficient refresh (every fifth and tenth
step), the F_TYPE 0 controller, the sponge with its CU2 zeroing, and the
warm start.
"""
import numpy as np
import pytest

from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import YMIN, ZMAX, ZMIN, config, mg_factors

pytestmark = pytest.mark.gpu


def _small(n: int):
    cfg = config(n)
    p = cfg.params
    if n != 1:
        p.nx, p.ny, p.nz = 32, 31, 31
        p.mg_ixp, p.mg_iex = mg_factors(p.nz + 2)
        p.mg_jyq, p.mg_jey = mg_factors(p.ny + 2)
    return cfg, p


def _case(name: str):
    if name == "config1":
        cfg, p = _small(1)
    elif name == "les_restart":
        cfg, p = _small(2)
        p.les_model = 2
    elif name == "f_type0":
        cfg, p = _small(1)
        p.f_type, p.wbulk0, p.px0 = 0, 0.3, 0.01
    elif name == "sponge_shipped":
        cfg, p = _small(2)
        for c in range(3):
            p.u_bc[c][ZMIN] = p.u_bc[c][ZMAX] = 6
        p.th_bc[YMIN] = 5
        p.th_bc[ZMIN] = p.th_bc[ZMAX] = 6
        p.mudpack_bc[0] = p.mudpack_bc[1] = 0
        p.f_type, p.wbulk0, p.px0 = 1, 0.05, 0.0
        p.sponge_type, p.ny_s_t, p.save_stats_int = 2, p.ny - 6, 2
        p.proj_mode = 0
    elif name == "headline_guess":
        cfg, p = _small(2)
        p.proj_mode, p.proj_guess = 2, 1
    else:
        raise ValueError(name)
    return cfg, p


@pytest.mark.parametrize("name", ["config1", "les_restart", "f_type0", "sponge_shipped", "headline_guess"])
def test_graph_replay_is_bit_identical(name):
    cfg, p = _case(name)
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    runs = []
    for graph in (False, True):
        s = S.DuctSolver(p, xp, yp)
        s.set_initial_physical(*ic)
        if name == "les_restart":
            # a restart at step 8: the first step refreshes in every substep,
            # step 10 refreshes both models at its first substep
            st = s.get_state()
            s.set_state(*st[:4], st[4], time=0.0, time_step=8)
        diags = []
        for _ in range(3 if name != "les_restart" else 4):
            for _q in range(3):
                s.rk_duct(graph=graph)
            d = s.diag()
            diags.append((d.div_ref, d.div_max, d.delta_t, d.time, d.time_step, d.mg_cycles, d.status))
        runs.append((s.get_state(), diags, s.forcing()))
        s.close()
    (sa, da, fa), (sb, db, fb) = runs
    for f, (a, b) in enumerate(zip(sa, sb)):
        assert np.array_equal(a, b), (name, f, np.abs(a - b).max())
    print(f"[graph {name}] diags {da[-1]}", flush=True)
    # the divergence diagnostic is summed with atomics, whose order differs
    # from run to run in either path; everything else is exact
    for x, y in zip(da, db):
        assert x[1:] == y[1:]
        assert abs(x[0] - y[0]) <= 1e-12 * abs(y[0])
    assert fa == fb
    assert all(d[-1] == 0 for d in da)


def test_sticky_status_from_the_device():
    """DELTA_T below 1e-4 (the reference's 'DELTA_T is zero' STOP,
    src/duct.F90:2812) is found by the projection's check on the device and
    reported by the initial projection, then by ekm_time_step and
    ekm_get_diag for both launch paths (it stays set for later substeps)."""
    cfg, p = _small(1)
    p.delta_t, p.variable_dt = 5.0e-5, 0
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    for graph in (False, True):
        s = S.DuctSolver(p, xp, yp)
        with pytest.raises(S.EkmError) as e:
            s.set_initial_physical(*ic)
        assert e.value.code == -5
        with pytest.raises(S.EkmError) as e:
            s.time_step(1, graph=graph)
        assert e.value.code == -5
        assert s.diag().status == -5
        s.close()
