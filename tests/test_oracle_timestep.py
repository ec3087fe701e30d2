"""Pins of the oracle's time-step functions by closed forms, invariants and
convergence orders -- not by agreement with the GPU.

Each test is chosen so that a plausible transcription mistake in the oracle
(a wrong RK weight, a flipped Coriolis sign, a missing metric term or
factor 1/2, a wrong face index) fails it:

- RK_DUCT's RKW3 constants and the Coriolis terms (src/diablo.F90:568-576,
  src/duct.F90:471-473): a uniform flow with nu = 0 oscillates about the
  geostrophic wind; one step must be exactly the third-order stability
  polynomial 1 + z + z^2/2 + z^3/6 of z = i f dt (any 3-stage 3rd-order RK),
  and a period takes 2 pi / f up to the scheme's O(dt^4) phase error.
- Crank-Nicolson ADI (src/duct.F90:1640-2391, THOMAS_REAL): on a uniform
  grid the discrete Dirichlet modes sin(m pi j/(NY+1)) decay per substep by
  (1 + nu h lambda/2)/(1 - nu h lambda/2) with the textbook eigenvalues
  lambda = -(4/dy^2) sin^2(m pi/(2(NY+1))).
- The laminar Ekman spiral is a steady solution of the continuous equations;
  the discrete step applied to it on tanh-stretched grids changes it by
  O(h^2) on rows J >= 2 (second-order convergence over three grids) and by
  O(h) on row 1, whose wall-face metric JACOBI_TRANS extrapolates linearly
  (src/diablo.F90:1193-1217).
- courant_curv_mpi (src/duct.F90:3605-3674): closed form for uniform flow.
- Advection and diffusion fluxes (src/duct.F90:683-1161, 1169-1263; the
  routines are shared by the four transported fields): the scalar's tendency
  on a sheared, stretched grid converges at second order to
  -div(u theta) + v + kappa lap(theta), the continuous right-hand side with
  the constant-stratification term.
- cofcr (src/duct.F90:3900-3934) through MUDPACK's discretisation: on a
  sheared grid the 9-point operator converges at second order to
  J lap(phi) - kx^2 J phi.
- Lagged periodic rows (src/boundary.F90:580-587 and the scalar's z rows):
  row 0 takes the entering field's NZ value, row NZ+1 its value at 1.
"""
import math

import numpy as np
import pytest

from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.params import YMAX, YMIN, ZMAX, ZMIN, base_params


def _params(nx, ny, nz, n_th=0):
    p = base_params(nx, ny, nz, n_th)
    p.variable_dt = 0
    p.f_type = 3
    p.f0 = 0.0
    p.nu = 0.0
    p.coriolis_mean_only = 1
    return p


def _zc(p):
    return np.zeros(p.shape_fourier, dtype=np.complex128)


def _to_fourier(p, f):
    """Physical (NY+2, NZ+2, NX+2) -> Fourier (NY+2, NZ+2, NX/2+1): the
    forward transform scaled by 1/NX with the modes above NX/3 zeroed."""
    c = np.fft.rfft(f[..., :p.nx], axis=-1) / p.nx
    c[..., p.nx // 3 + 1:] = 0.0
    return np.ascontiguousarray(c)


def _rk_hbar(dt):
    return np.array([8.0, 2.0, 5.0]) / 15.0 * dt


def _centres(xp, yp):
    """(z, y) of the unknowns.  GRID_YZ.dat lists cell vertices and the
    metrics of JACOBI_TRANS are those of the cell spanned by vertices K..K+1,
    J..J+1 (its faces at vertex row J, J+1 and column K, K+1,
    src/diablo.F90:1166-1190, 1268-1292), so unknown (K, J) sits at that
    cell's centre; the last row and column are extrapolated linearly."""
    def c(a):
        m = 0.25 * (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:])
        m = np.concatenate([m, 2 * m[:, -1:] - m[:, -2:-1]], axis=1)
        return np.concatenate([m, 2 * m[-1:, :] - m[-2:-1, :]], axis=0)
    return c(xp), c(yp)


# ---------------------------------------------------------------------------
def test_inertial_oscillation_is_the_rk3_polynomial(oracle_lib):
    """A uniform (u, w) about zero geostrophic wind, nu = 0, stress-free
    ends: du/dt = -f w, dw/dt = f u.  The amplitude is small so that the
    advection terms (quadratic; nonzero in the boundary cells, whose rows
    the reference fixes, e.g. U1 = 0 on the K = NZ+1 row at src/boundary.F90
    :792-797) stay below 1e-9 of the rotation."""
    O = oracle_lib
    p = _params(8, 7, 7)
    p.f0, p.delta_t = 0.3, 0.7
    for s in (YMIN, YMAX):            # u, w stress-free, v = 0 at both y ends
        p.u_bc[0][s], p.u_bc[1][s], p.u_bc[2][s] = 1, 0, 1
    for s in (ZMIN, ZMAX):
        p.u_bc[0][s], p.u_bc[1][s], p.u_bc[2][s] = 1, 1, 1
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 2, 2, 2, 1
    xp, yp = tanh_grid(p.ny, p.nz, 1.0, 1.0, 0.0)
    o = O.Oracle(p, xp, yp)
    amp = 1e-11
    u0, w0 = 0.7 * amp, -0.4 * amp
    cu1, cu3 = _zc(p), _zc(p)
    cu1[..., 0], cu3[..., 0] = u0, w0
    o.set_state(cu1, _zc(p), cu3, _zc(p))
    z = 1j * p.f0 * p.delta_t
    P = 1 + z + z ** 2 / 2 + z ** 3 / 6
    c0 = u0 + 1j * w0
    inner = (slice(2, p.ny), slice(2, p.nz))
    n_period = int(round(2 * math.pi / (p.f0 * p.delta_t)))
    worst = 0.0
    for n in range(1, n_period + 1):
        assert o.time_step(1) == 0
        cu1, cu2, cu3, cp, _ = o.get_state()
        c = cu1[inner + (0,)].real + 1j * cu3[inner + (0,)].real
        worst = max(worst, np.abs(c - c0 * P ** n).max() / abs(c0))
        assert np.abs(cu1[..., 1:]).max() < 1e-20
    print(f"inertial oscillation: worst relative deviation from P(z)^n {worst:.2e}")
    assert worst < 1e-9
    # a mistyped RK weight changes P at O(z^2) or O(z^3); the closest
    # alternatives (Euler's 1 + z, Heun's 1 + z + z^2/2) are 1e-3 away
    assert abs(P - (1 + z + z ** 2 / 2)) > 1e-3
    # one period 2 pi / f: back to the start up to the scheme's phase and
    # amplitude errors, n (f dt)^4 / 24 to leading order
    err = np.abs(c - c0 * np.exp(1j * p.f0 * p.delta_t * n_period)).max() / abs(c0)
    bound = n_period * (p.f0 * p.delta_t) ** 4 / 24
    assert err < 1.5 * bound + abs(2 * math.pi - n_period * p.f0 * p.delta_t)


def test_crank_nicolson_decay_of_discrete_eigenmodes(oracle_lib):
    """W (U3) with Dirichlet rows at J = 0 and J = NY, NY+1 (the reference
    sets both top rows of U2 and U3, src/boundary.F90:1410-1419) and
    stress-free z ends, z-uniform: the unknowns J = 1..NY-1 of a uniform y
    grid carry the modes sin(m pi J / NY)."""
    O = oracle_lib
    p = _params(8, 15, 7)
    p.nu, p.delta_t = 0.05, 0.3
    ly = 2.0
    for c in range(3):
        p.u_bc[c][YMIN] = p.u_bc[c][YMAX] = 0
        p.u_bc_c1[c][YMIN] = p.u_bc_c1[c][YMAX] = 0.0
    for s in (ZMIN, ZMAX):
        p.u_bc[0][s], p.u_bc[1][s], p.u_bc[2][s] = 1, 1, 1
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 2, 2, 2, 1
    xp, yp = tanh_grid(p.ny, p.nz, ly, 1.0, 0.0)
    o = O.Oracle(p, xp, yp)
    n = p.ny
    dy = ly / (p.ny + 1)
    j = np.arange(p.ny + 2)
    modes = {1: 0.8, 4: 0.3, 11: 0.05}

    def shape(m):
        s_ = np.sin(m * np.pi * j / n)
        s_[n:] = 0.0
        return s_

    w = sum(a * shape(m) for m, a in modes.items())
    cu3 = _zc(p)
    cu3[..., 0] = w[:, None]
    o.set_state(_zc(p), _zc(p), cu3, _zc(p))
    assert o.time_step(1) == 0
    cu1, cu2, cu3, cp, _ = o.get_state()
    expect = np.zeros_like(w)
    for m, a in modes.items():
        lam = -(4.0 / dy ** 2) * np.sin(m * np.pi / (2 * n)) ** 2
        g = np.prod([(1 + p.nu * h * lam / 2) / (1 - p.nu * h * lam / 2) for h in _rk_hbar(p.delta_t)])
        expect = expect + a * g * shape(m)
    # the reference re-solves row J = NY in U3's z sweep (JEND = NY for
    # W_BC_YMAX = 0, src/duct.F90:102-104, :2329), after the y sweep set it
    # to C1; that row is not part of the y modes
    err = np.abs(cu3[:n, :, 0].real - expect[:n, None]).max()
    print(f"CN eigenmodes: max deviation {err:.2e}")
    assert err < 1e-14
    assert max(np.abs(cu1).max(), np.abs(cu2).max(), np.abs(cp).max()) < 1e-14
    # the modes decay at different rates (mode 11 by about 1e-3 per step)
    assert np.abs(expect).max() > 0.5


def _ekman_tendency(O, n_y):
    p = _params(8, n_y - 2, 7)
    re = 500.0
    p.nu, p.f0, p.delta_t = 1.0 / re, 2.0 / re, 1e-4     # Ekman depth sqrt(2 nu / f) = 1
    ug = 1.0
    for c in range(3):
        p.u_bc[c][YMIN] = 0
        p.u_bc_c1[c][YMIN] = 0.0
    p.u_bc[0][YMAX], p.u_bc_c1[0][YMAX] = 0, ug
    p.u_bc[1][YMAX], p.u_bc_c1[1][YMAX] = 0, 0.0
    p.u_bc[2][YMAX], p.u_bc_c1[2][YMAX] = 1, 0.0
    for c in range(3):
        p.u_bc[c][ZMIN] = p.u_bc[c][ZMAX] = 6
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 0, 0, 2, 1
    xp, yp = tanh_grid(p.ny, p.nz, 20.0, 1.0, 2.0)
    yc = _centres(xp, yp)[1][:, 0]
    y = yc - yc[0]            # the Dirichlet row 0 is the wall
    u = ug * (1.0 - np.exp(-y) * np.cos(y))
    w = -ug * np.exp(-y) * np.sin(y)
    cu1, cu3 = _zc(p), _zc(p)
    cu1[..., 0], cu3[..., 0] = u[:, None], w[:, None]
    o = O.Oracle(p, xp, yp)
    o.set_state(cu1, _zc(p), cu3, _zc(p))
    assert o.time_step(1) == 0
    st = o.get_state()
    d = np.maximum(np.abs(st[0][..., 0].real - u[:, None]), np.abs(st[2][..., 0].real - w[:, None])) / p.delta_t
    ks = slice(1, p.nz + 1)
    return d[1, ks].max(), d[2:p.ny + 1, ks].max()


def test_laminar_ekman_spiral_is_steady_to_second_order(oracle_lib):
    r = [_ekman_tendency(oracle_lib, n) for n in (65, 129, 257)]
    e1, e = [a for a, _ in r], [b for _, b in r]
    print("Ekman spiral tendency per grid: row 1", e1, "rows >= 2", e)
    assert 3.3 < e[0] / e[1] < 4.7 and 3.3 < e[1] / e[2] < 4.7, e
    assert 1.7 < e1[0] / e1[1] < 2.3 and 1.7 < e1[1] / e1[2] < 2.3, e1
    assert e[2] < 1e-6 and e1[2] < 1e-4


@pytest.mark.parametrize("u0,w0", [(0.9, 0.05), (0.02, -0.6)])
def test_cfl_time_step_of_a_uniform_flow(u0, w0, oracle_lib):
    O = oracle_lib
    p = _params(16, 15, 15)
    p.variable_dt, p.cfl, p.delta_t = 1, 0.5, 10.0
    for s in (YMIN, YMAX):
        p.u_bc[0][s], p.u_bc[1][s], p.u_bc[2][s] = 1, 0, 1
    for c in range(3):
        p.u_bc[c][ZMIN] = p.u_bc[c][ZMAX] = 6
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 0, 0, 2, 1
    lx, ly, lz = 3.0, 2.0, 1.5
    p.lx = lx
    # uniform grid: the cell Jacobian (an average of face Jacobians) is then
    # exactly dz dy, so the bound is exactly CFL dz / |w|
    xp, yp = tanh_grid(p.ny, p.nz, ly, lz, 0.0)
    o = O.Oracle(p, xp, yp)
    cu1, cu3 = _zc(p), _zc(p)
    cu1[..., 0], cu3[..., 0] = u0, w0
    o.set_state(cu1, _zc(p), cu3, _zc(p))
    assert o.time_step(1) == 0
    dz = lz / (p.nz + 1)
    expect = min(p.cfl * (lx / p.nx) / abs(u0), p.cfl * dz / abs(w0))
    assert abs(o.diag().delta_t - expect) <= 1e-14 * expect
    # capped at DELTA_T_SET
    p2 = p.copy()
    p2.delta_t = 0.5 * expect
    o2 = O.Oracle(p2, xp, yp)
    o2.set_state(cu1, _zc(p), cu3, _zc(p))
    assert o2.time_step(1) == 0
    assert o2.diag().delta_t == p2.delta_t


# ---------------------------------------------------------------------------
def _sheared_grid(p, ly, lz, shear, gamma):
    """z of every (K, J) sheared along y: z = K lz/(NZ+1) + shear sin(pi y / ly),
    y tanh-stretched."""
    xp, yp = tanh_grid(p.ny, p.nz, ly, lz, gamma)
    return xp + shear * np.sin(np.pi * yp / ly), yp


def _transport_error(O, n):
    p = _params(16, n - 2, n - 2, n_th=1)
    kappa = 0.05
    p.nu, p.pr, p.cont_strat, p.ri_tau = kappa, 1.0, 1, 0.0
    p.delta_t = 1e-4
    lx, ly, lz = 2 * np.pi, 2.0, 2.0
    p.lx, p.ly, p.lz = lx, ly, lz
    for s in range(4):
        p.th_bc[s], p.th_bc_c1[s] = 1, 0.0
    xp, yp = _sheared_grid(p, ly, lz, 0.15, 1.2)
    o = O.Oracle(p, xp, yp)
    zc, yc = _centres(xp, yp)
    x = np.arange(p.nx + 2) * lx / p.nx
    X, Y, Z = x[None, None, :], yc[:, :, None], zc[:, :, None]
    a = 2 * np.pi / lx
    u = 0.3 + 0.2 * np.cos(a * X) * np.sin(Y) + 0 * Z
    v = 0.25 * np.sin(np.pi * Y / ly) * (1 + 0.3 * np.cos(Z)) + 0 * X
    w = 0.2 + 0.15 * np.sin(Z) * np.cos(Y) + 0 * X
    th = 0.5 * np.cos(Y) + 0.2 * np.sin(a * X) * np.sin(Z) * np.cos(0.7 * Y)
    # d(u th)/dx, d(v th)/dy, d(w th)/dz and lap(th), by hand
    sX, cX, sY, cY, sZ, cZ = np.sin(a * X), np.cos(a * X), np.sin(Y), np.cos(Y), np.sin(Z), np.cos(Z)
    c7, s7 = np.cos(0.7 * Y), np.sin(0.7 * Y)
    u_x = -0.2 * a * sX * sY
    th_x = 0.2 * a * cX * sZ * c7
    th_y = -0.5 * sY - 0.2 * 0.7 * sX * sZ * s7
    th_z = 0.2 * sX * cZ * c7
    v_y = 0.25 * (np.pi / ly) * np.cos(np.pi * Y / ly) * (1 + 0.3 * cZ)
    w_z = 0.15 * cZ * cY
    lap = (-0.2 * a * a * sX * sZ * c7) + (-0.5 * cY - 0.2 * 0.49 * sX * sZ * c7) + (-0.2 * sX * sZ * c7)
    rhs = -(u_x * th + u * th_x) - (v_y * th + v * th_y) - (w_z * th + w * th_z) + v + kappa * lap
    o.set_state(_to_fourier(p, u), _to_fourier(p, v), _to_fourier(p, w), _zc(p), _to_fourier(p, th))
    th0 = o.get_state()[4]
    assert o.rk_duct() == 0
    th1 = o.get_state()[4]
    h = _rk_hbar(p.delta_t)[0]
    got = np.fft.irfft((th1 - th0) / h * p.nx, n=p.nx, axis=-1)
    # rows next to the boundary rows (which the scalar's Neumann rows set)
    # feel them through the implicit sweeps; two rows further in that is
    # below the O(h^2) error
    sl = (slice(4, p.ny - 2), slice(4, p.nz - 2), slice(0, p.nx))
    return np.abs(got[sl] - rhs[sl]).max() / np.abs(rhs[sl]).max()


def test_scalar_transport_is_second_order_on_a_sheared_grid(oracle_lib):
    e = [_transport_error(oracle_lib, n) for n in (33, 65, 129)]
    print("scalar transport relative error per grid:", e)
    assert 3.3 < e[0] / e[1] < 4.8 and 3.3 < e[1] / e[2] < 4.8, e
    assert e[2] < 2e-3


def _apply9(cf, phi):
    """MUDPACK's 9-point stencil (c1 E, c2 NE, c3 N, c4 NW, c5 W, c6 SW,
    c7 S, c8 SE, c9 centre) on the interior of phi (ny, nx)."""
    c = phi[1:-1, 1:-1]
    E, W = phi[1:-1, 2:], phi[1:-1, :-2]
    N, S = phi[2:, 1:-1], phi[:-2, 1:-1]
    NE, NW, SW, SE = phi[2:, 2:], phi[2:, :-2], phi[:-2, :-2], phi[:-2, 2:]
    k = [a[1:-1, 1:-1] for a in cf[:9]]
    return k[0] * E + k[1] * NE + k[2] * N + k[3] * NW + k[4] * W + k[5] * SW + k[6] * S + k[7] * SE + k[8] * c


def _cofcr_error(O, n, kx2):
    p = base_params(8, n - 2, n - 2)
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 2, 2, 2, 1
    xp, yp = _sheared_grid(p, 2.0, 2.0, 0.2, 1.5)
    o = O.Oracle(p, xp, yp)
    cf = o.mg_discretize(kx2, o.mg_ngrid())
    z, y = _centres(xp, yp)
    phi = np.cos(1.3 * z) * np.sin(0.7 * y + 0.2) + 0.2 * z * y
    lap = -(1.3 ** 2 + 0.7 ** 2) * np.cos(1.3 * z) * np.sin(0.7 * y + 0.2)
    jac = o.metric(7)[: p.ny + 2, : p.nz + 2]
    expect = (jac * (lap - kx2 * phi))[1:-1, 1:-1]
    got = _apply9(cf, phi)
    sl = (slice(2, -2), slice(2, -2))
    return np.abs(got[sl] - expect[sl]).max() / np.abs(expect[sl]).max()


@pytest.mark.parametrize("kx2", [0.0, 2.5])
def test_cofcr_operator_is_second_order_on_a_sheared_grid(kx2, oracle_lib):
    """Above about 129 points per direction the metrics' truncation to 1e-8
    (src/diablo.F90:1220-1230), differenced by cx and cy, sets a floor near
    5e-5 on this grid, so the order is measured on two pairs below it."""
    e = [_cofcr_error(oracle_lib, n, kx2) for n in (41, 81, 65, 129)]
    print(f"cofcr operator relative error per grid (kx2={kx2}):", e)
    assert 3.3 < e[0] / e[1] < 4.8 and 3.3 < e[2] / e[3] < 4.8, e
    assert e[3] < 1e-4


def test_lagged_periodic_rows(oracle_lib):
    """Scalar with type-6 z rows, at rest and without diffusion: the ADI is
    the identity except the boundary rows, which take the entering field's
    values at K = NZ (row 0) and K = 1 (row NZ+1)."""
    O = oracle_lib
    p = _params(8, 15, 15, n_th=1)
    p.cont_strat, p.pr, p.delta_t = 1, 1.0, 0.1
    for c in range(3):
        p.u_bc[c][YMIN] = p.u_bc[c][YMAX] = 0
        p.u_bc[c][ZMIN] = p.u_bc[c][ZMAX] = 6
    p.th_bc[YMIN] = p.th_bc[YMAX] = 1
    p.th_bc[ZMIN] = p.th_bc[ZMAX] = 6
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = 0, 0, 2, 1
    xp, yp = tanh_grid(p.ny, p.nz, 1.0, 1.0, 0.0)
    o = O.Oracle(p, xp, yp)
    rng = np.random.default_rng(5)
    cth = (rng.standard_normal(p.shape_fourier) + 1j * rng.standard_normal(p.shape_fourier))
    cth[..., p.nx // 3 + 1:] = 0.0
    cth[..., 0] = cth[..., 0].real
    cth[:, 0, :] = 7.0       # entering ghost values that must be overwritten
    cth[:, -1, :] = -7.0
    o.set_state(_zc(p), _zc(p), _zc(p), _zc(p), np.ascontiguousarray(cth))
    assert o.rk_duct() == 0
    th1 = o.get_state()[4]
    rows = slice(2, p.ny)
    tol = 1e-14 * np.abs(cth).max()     # one physical <-> Fourier round trip
    assert np.abs(th1[rows, 0, :] - cth[rows, p.nz, :]).max() < tol
    assert np.abs(th1[rows, p.nz + 1, :] - cth[rows, 1, :]).max() < tol
    assert np.abs(th1[rows, 1:p.nz + 1, :] - cth[rows, 1:p.nz + 1, :]).max() < tol
    assert abs(cth[5, 0, 1] - th1[5, 0, 1]) > 1.0    # the entering ghost value was replaced
