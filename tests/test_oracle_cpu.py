"""CPU checks that pin the oracle (no GPU needed).

Each test checks one oracle function against something independent of the
oracle's own formulas: the DFT definition (numpy), dense linear algebra
(numpy / scipy), closed forms on rectilinear grids, the reference's own
MUDPACK compiled from src/mud2cr.f and src/mudcom.f (oracle/_ref, skipped
where it is not built), the reference's shipped grid (tests/golden), and the
projection identity on a uniform grid.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import base_params, config

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


# -- A3: x-FFT ---------------------------------------------------------------
@pytest.mark.parametrize("nx", [8, 64, 256, 1024])
def test_fft_is_scaled_truncated_dft(nx, oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(nx)
    v = rng.standard_normal((11, nx + 2))
    v[:, nx:] = 0.0
    a = v.copy()
    O.fft_fwd(a)
    ref = np.fft.rfft(v[:, :nx], axis=1) / nx
    ref[:, nx // 3 + 1:] = 0.0
    got = a.view(np.complex128)
    assert np.abs(got - ref).max() < 1e-14 * max(1.0, np.abs(ref).max()) * math.log2(nx)
    # inverse: unscaled synthesis from modes 0..NX/3 with Hermitian symmetry
    O.fft_inv(a)
    back = np.fft.irfft(ref * nx, n=nx, axis=1)
    assert np.abs(a[:, :nx] - back).max() < 1e-13
    assert np.all(a[:, nx:] == 0.0)


def test_fft_derivative_of_sine(oracle_lib):
    """d/dx of sin(m x) through i*kx in Fourier space is m cos(m x)."""
    O = oracle_lib
    nx, lx, m = 64, 2 * math.pi, 5
    x = np.arange(nx) * lx / nx
    v = np.zeros((1, nx + 2))
    v[0, :nx] = np.sin(m * x)
    O.fft_fwd(v)
    c = v.view(np.complex128)[0]
    kx = 2 * math.pi * np.arange(nx // 2 + 1) / lx
    c *= 1j * kx
    O.fft_inv(v)
    assert np.abs(v[0, :nx] - m * np.cos(m * x)).max() < 1e-12


# -- A6: Thomas ----------------------------------------------------------------
def _dense_solve(a, b, c, g, extra=None):
    """Solve every batch column of the tridiagonal systems densely."""
    ny1, nb = g.shape
    out = np.zeros_like(g)
    for i in range(nb):
        M = np.diag(b[:, i]) + np.diag(a[1:, i], -1) + np.diag(c[:-1, i], 1)
        if extra is not None:
            r, col, d = extra
            M[r, col] += d
        out[:, i] = np.linalg.solve(M, g[:, i])
    return out


def test_thomas_matches_dense(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(11)
    ny, nx = 30, 17
    a = rng.uniform(-1, 0, (ny + 1, nx + 1))
    c = rng.uniform(-1, 0, (ny + 1, nx + 1))
    b = 2.5 + rng.uniform(0, 1, (ny + 1, nx + 1))
    g = rng.standard_normal((ny + 1, nx + 1))
    ref = _dense_solve(a, b, c, g)
    x = g.copy()
    O.thomas_real(a.copy(), b.copy(), c.copy(), x)
    assert rel(x, ref) < 1e-13


def test_thomas_sp_matches_dense(oracle_lib):
    """THOMAS_REAL_SP eliminates row K+1's extra entry D at column K-1 with
    row K (src/boundary.F90:17-24) before the plain sweep."""
    O = oracle_lib
    rng = np.random.default_rng(12)
    ny, nx, k, d = 24, 9, 20, 0.37
    a = rng.uniform(-1, -0.2, (ny + 1, nx + 1))
    c = rng.uniform(-1, 0, (ny + 1, nx + 1))
    b = 3.0 + rng.uniform(0, 1, (ny + 1, nx + 1))
    g = rng.standard_normal((ny + 1, nx + 1))
    ref = _dense_solve(a, b, c, g, extra=(k + 1, k - 1, d))
    x = g.copy()
    O.thomas_real_sp(a.copy(), b.copy(), c.copy(), x, d, k)
    assert rel(x, ref) < 1e-12


# -- A13: metrics ----------------------------------------------------------------
def _oracle_for_grid(O, ny, nz, xp, yp, bcs=(2, 2, 2, 2)):
    p = base_params(16, ny, nz)
    for s in range(4):
        p.mudpack_bc[s] = bcs[s]
    return O.Oracle(p, xp, yp), p


def test_metrics_uniform_rectilinear_closed_form(oracle_lib):
    O = oracle_lib
    ny, nz, dy, dz = 15, 31, 0.5, 0.25
    y = dy * np.arange(ny + 2)
    z = dz * np.arange(nz + 2)
    xp, yp = np.meshgrid(z, y)
    o, _ = _oracle_for_grid(O, ny, nz, xp, yp)
    sl = (slice(None), slice(0, ny + 2), slice(0, nz + 2))
    g11, g12, g22 = o.metric(0)[sl], o.metric(1)[sl], o.metric(2)[sl]
    c11, c12, c21, c22 = (o.metric(w)[sl] for w in (3, 4, 5, 6))
    jac = o.metric(7)[:ny + 2, :nz + 2]
    # interior faces (K,J = 1..NZ+1, 1..NY+1) and the extrapolated ghost rows
    inner = (slice(None), slice(1, ny + 2), slice(1, nz + 2))
    assert np.all(g11[inner] == dy / dz) and np.all(g22[inner] == dz / dy)
    assert np.all(g12[inner] == 0.0) and np.all(c12[inner] == 0.0) and np.all(c21[inner] == 0.0)
    assert np.all(c11[inner] == dy) and np.all(c22[inner] == dz)
    assert np.all(jac == dy * dz)


def test_metrics_truncate_to_1e8(oracle_lib):
    """x_zeta etc. are cut to 8 decimals before the metrics are formed
    (src/diablo.F90:1220-1230): CJOB_22 = x_zeta carries the cut value."""
    O = oracle_lib
    ny, nz, dy, dz = 7, 15, 0.5, 0.123456789123
    xp, yp = np.meshgrid(dz * np.arange(nz + 2), dy * np.arange(ny + 2))
    o, _ = _oracle_for_grid(O, ny, nz, xp, yp)
    c22 = o.metric(6)[:, 2:ny, 2:nz]
    assert np.abs(c22 - 0.12345678).max() < 1e-15
    assert np.abs(o.metric(7)[2:ny, 2:nz] - 0.12345678 * 0.5).max() < 1e-15


def test_metrics_stretched_rectilinear(oracle_lib):
    """On a tanh-stretched y, uniform z grid the metrics reduce to face
    spacings: y_eta on the zeta face is y(j+1)-y(j), on the eta face the
    centred (y(j+1)-y(j-1))/2 (src/diablo.F90:1166-1190)."""
    O = oracle_lib
    ny, nz = 31, 15
    xp, yp = tanh_grid(ny, nz, 20.0, 2 * math.pi, 2.0)
    o, _ = _oracle_for_grid(O, ny, nz, xp, yp)
    y, z = yp[:, 0], xp[0, :]
    t8 = lambda v: np.trunc(v * 1e8) * 1e-8  # noqa: E731
    dz = t8(z[1] - z[0])
    c11 = o.metric(3)          # CJOB_11 = y_eta
    c22 = o.metric(6)          # CJOB_22 = x_zeta
    for j in range(2, ny):
        for k in range(2, nz):
            ye1 = t8(0.25 * (2 * y[j + 1] - 2 * y[j - 1]))
            ye2 = t8(y[j + 1] - y[j])
            assert abs(c11[0, j, k] - ye1) <= 1e-15 * abs(ye1)
            assert abs(c11[1, j, k] - ye2) <= 1e-15 * abs(ye2)
            assert abs(c22[0, j, k] - dz) <= 1e-15
    g12 = o.metric(1)[:, 1:ny + 2, 1:nz + 2]
    assert np.all(g12 == 0.0)


def test_les_metric_variants_rectilinear_closed_form(oracle_lib):
    """Subgrid-stress metrics (src/diablo.F90:1379-1474) on a uniform
    rectilinear grid: the U3 stress doubles the zeta_x products and the U2
    stress the eta_y products, so GMAT_11_z = 2 GMAT_11, GMAT_11_y = GMAT_11,
    GMAT_22_z = GMAT_22, GMAT_22_y = 2 GMAT_22 and both GMAT_12 variants are 0
    on every face, ghosts included."""
    O = oracle_lib
    ny, nz, dy, dz = 15, 31, 0.5, 0.25
    xp, yp = np.meshgrid(dz * np.arange(nz + 2), dy * np.arange(ny + 2))
    o, _ = _oracle_for_grid(O, ny, nz, xp, yp)
    g11, g22 = o.metric(0), o.metric(2)
    g11z, g12z, g22z, g11y, g12y, g22y = (o.metric(w) for w in range(8, 14))
    assert np.all(g11 == dy / dz) and np.all(g22 == dz / dy)
    assert np.array_equal(g11z, 2 * g11) and np.array_equal(g11y, g11)
    assert np.array_equal(g22z, g22) and np.array_equal(g22y, 2 * g22)
    assert np.all(g12z == 0.0) and np.all(g12y == 0.0)


def test_les_metric_variants_sum_to_three_times_molecular(oracle_lib):
    """On a skewed grid every metric is nonzero; the _z and _y variants put
    the factor 2 on complementary products, so X_z + X_y = 3 X for X = GMAT_11,
    GMAT_12, GMAT_22 on every face, including the extrapolated ghosts and the
    corner rules (which treat the variants as the reference treats X)."""
    O = oracle_lib
    ny, nz = 31, 15
    xp, yp = tanh_grid(ny, nz, 20.0, 2 * math.pi, 2.0)
    xp, yp = xp + 0.05 * np.sin(0.3 * yp), yp + 0.02 * np.sin(0.5 * xp)
    o, _ = _oracle_for_grid(O, ny, nz, xp, yp)
    for base, vz, vy in ((0, 8, 11), (1, 9, 12), (2, 10, 13)):
        g, gz, gy = o.metric(base), o.metric(vz), o.metric(vy)
        fin = np.isfinite(g)
        assert np.array_equal(fin, np.isfinite(gz)) and np.array_equal(fin, np.isfinite(gy))
        assert np.abs(g[fin]).max() > 0.0
        err = np.abs(gz[fin] + gy[fin] - 3.0 * g[fin]).max()
        assert err <= 1e-12 * np.abs(g[fin]).max(), (base, err)
        assert np.abs(gz[fin] - g[fin]).max() > 1e-6 * np.abs(g[fin]).max()
        assert np.abs(gy[fin] - g[fin]).max() > 1e-6 * np.abs(g[fin]).max()


def test_les_adi_constant_viscosity_on_u1_is_plain_adi(oracle_lib):
    """U1's subgrid terms use the molecular metrics GMAT_11 / GMAT_22
    (src/duct.F90:1667-1731, :1771-1832), so with a constant eddy viscosity
    c the ADI with weight t2 equals the plain ADI with t1 + t2 c: the
    explicit pre-terms, the matrix rows and the emitted right-hand side all
    scale the same face averages.  A varying c(i) along x does the same line
    by line.  Skewed config 1 grid, random U and R."""
    from oracle import oracle as O
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    xp = xp + 0.05 * np.sin(0.3 * yp)
    o = O.Oracle(p, xp, yp)
    rng = np.random.default_rng(7)
    shape = (p.ny + 2, p.nz + 2, p.nx + 2)
    u = rng.standard_normal(shape)
    r = rng.standard_normal(shape)
    t1, t2 = 2.0e-3, 0.05
    ci = 1.0e-2 * (1.0 + rng.random(p.nx + 2))
    for c in (np.full(p.nx + 2, 3.0e-2), ci):
        nut = np.broadcast_to(c, shape).copy()
        ul, rl = o.adi_field(0, u, r, t1, nut, t2)
        for i in (0, 5, p.nx):
            up, _ = o.adi_field(0, u, r, t1 + t2 * c[i])
            err = np.abs(ul[..., i] - up[..., i]).max()
            assert err <= 1e-12 * np.abs(up[..., i]).max(), (i, err)
    # the viscosity changes the solution (the check is not vacuous)
    up, _ = o.adi_field(0, u, r, t1)
    assert np.abs(ul - up).max() > 1e-3 * np.abs(up).max()


def test_shipped_grid_is_tanh_family():
    """rundir_unst/GRID_YZ.dat (fixture) is the tanh family with gamma 2.3:
    y uniform in z, z uniform, to the file's printed precision."""
    d = json.load(open(os.path.join(GOLDEN, "grid_yz_unst.json")))
    g = d["grid_def"]
    ny, nz = g["ny"], g["nz"]
    y, z = np.array(d["y"]), np.array(d["z"])
    xp, yp = tanh_grid(ny, nz, y[-1], z[-1], 2.3)
    assert np.abs(yp[:, 0] - y).max() < 1e-12
    assert np.abs(xp[0, :] - z).max() < 1e-12
    for j, k, v in d["y_exceptions"]:
        assert abs(v - y[j]) < 1e-12


# -- A10/A11: multigrid --------------------------------------------------------
@pytest.mark.parametrize("bcs,kx2", [((2, 2, 2, 1), 0.0), ((2, 2, 2, 2), 0.25), ((0, 0, 2, 2), 4.0),
                                     ((1, 1, 1, 1), 0.0), ((2, 2, 0, 0), 1.0)])
def test_mudpack_matches_reference_build(bcs, kx2, oracle_lib):
    """The oracle's MUDPACK restatement against mud2cr itself (compiled from
    the reference's src/mud2cr.f, src/mudcom.f): bit-for-bit, equal cycles."""
    O = oracle_lib
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference tree and amdflang)")
    cfg = config(1)
    p = cfg.params
    for s in range(4):
        p.mudpack_bc[s] = bcs[s]
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    xp = xp + 0.05 * np.sin(0.3 * yp)       # skewed: cross metrics nonzero
    o = O.Oracle(p, xp, yp)
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((p.ny + 2, p.nz + 2))
    phi_o, cyc_o = o.mudpack(kx2, rhs)
    phi_r, cyc_r, err = O.ref_mudpack(o, kx2, rhs)
    # ierror -1 after intl=1 is MUDPACK's "maxcy reached before tolmax" warning
    assert err[0] == 0 and err[1] in (0, -1)
    assert (err[1] == -1) == (cyc_r == p.mg_maxcy)
    assert cyc_o == cyc_r
    assert np.array_equal(phi_o, phi_r)


def _assemble(cf, nx, ny, bcs):
    """Sparse matrix of MUDPACK's finest-level 9-point system, one row per
    grid point (i fastest); Dirichlet rows are identity rows.  In a periodic
    direction every one of the n points carries its own equation and the
    virtual neighbours are phi(0) = phi(n-1), phi(n+1) = phi(2)
    (src/mud2cr.f:1693-1795)."""
    off = {0: (1, 0), 1: (1, 1), 2: (0, 1), 3: (-1, 1), 4: (-1, 0), 5: (-1, -1), 6: (0, -1), 7: (1, -1)}
    px, py = bcs[0] == 0, bcs[2] == 0
    mx, my = nx, ny

    def wrap(v, n, per):
        if per and v == -1:
            return n - 2
        if per and v == n:
            return 1
        return v

    idx = lambda i, j: wrap(i, nx, px) + mx * wrap(j, ny, py)  # noqa: E731
    rows, cols, vals = [], [], []
    for j in range(my):
        for i in range(mx):
            r = idx(i, j)
            spec = (bcs[0] == 1 and i == 0) or (bcs[1] == 1 and i == nx - 1) or \
                   (bcs[2] == 1 and j == 0) or (bcs[3] == 1 and j == ny - 1)
            if spec:
                rows.append(r); cols.append(r); vals.append(1.0)
                continue
            rows.append(r); cols.append(r); vals.append(cf[8, j, i])
            for q, (di, dj) in off.items():
                ii, jj = i + di, j + dj
                inside = (px or 0 <= ii < nx) and (py or 0 <= jj < ny)
                if not inside:
                    assert cf[q, j, i] == 0.0, (q, i, j)
                    continue
                spec_nb = (not px and ((bcs[0] == 1 and ii == 0) or (bcs[1] == 1 and ii == nx - 1))) or \
                          (not py and ((bcs[2] == 1 and jj == 0) or (bcs[3] == 1 and jj == ny - 1)))
                if spec_nb:
                    continue            # phi = 0 there
                rows.append(r); cols.append(idx(ii, jj)); vals.append(cf[q, j, i])
    n = mx * my
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n)), mx, my


@pytest.mark.parametrize("bcs,kx2", [((2, 2, 2, 1), 0.0), ((1, 1, 1, 1), 0.0), ((2, 2, 2, 2), 0.5),
                                     ((0, 0, 2, 1), 0.0), ((2, 2, 0, 0), 0.3)])
def test_mudpack_converges_to_dense_solution(bcs, kx2, oracle_lib):
    """With a tight relmax tolerance the W(2,1) cycles converge to the
    solution of the assembled finest-level 9-point system (sparse direct
    solve), boundary folding included.  The grid is isotropic in index space
    (h_y ~ h_z) and skewed so the cross terms are active: point relaxation
    converges fast there (on strongly stretched grids it does not, which is
    the reference's own behaviour)."""
    O = oracle_lib
    p = base_params(16, 63, 63)
    for s in range(4):
        p.mudpack_bc[s] = bcs[s]
    p.mg_maxcy, p.mg_tolmax = 60, 1e-12
    xp, yp = tanh_grid(p.ny, p.nz, 2 * math.pi, 2 * math.pi, 0.0)
    xp = xp + 0.05 * np.sin(yp)
    o = O.Oracle(p, xp, yp)
    cf = o.mg_discretize(kx2, o.mg_ngrid())
    ny, nx = cf.shape[1:]
    A, mx, my = _assemble(cf, nx, ny, bcs)
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((ny, nx))
    if bcs[0] == 0:
        rhs[:, nx - 1] = rhs[:, 0]      # periodic: point nx is point 1
    if bcs[2] == 0:
        rhs[ny - 1, :] = rhs[0, :]
    phi, cyc = o.mudpack(kx2, rhs)
    b = np.zeros(mx * my)
    for j in range(my):
        for i in range(mx):
            spec = (bcs[0] == 1 and i == 0) or (bcs[1] == 1 and i == nx - 1) or \
                   (bcs[2] == 1 and j == 0) or (bcs[3] == 1 and j == ny - 1)
            b[i + mx * j] = 0.0 if spec else -rhs[j, i]
    x = spla.spsolve(A.tocsc(), b).reshape(my, mx)
    assert cyc < 60
    assert rel(phi[:my, :mx], x) < 1e-9


def test_point_relaxation_stalls_on_stretched_grid(oracle_lib):
    """Documents the reference configuration's convergence on the stretched
    64x65x65 grid: the residual of the assembled system falls by only ~0.85
    per W(2,1) cycle (index-space anisotropy up to ~40 near the top), so six
    cycles end far from the discrete solution."""
    O = oracle_lib
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    res = []
    for cyc in (8, 16):
        p.mg_maxcy, p.mg_tolmax = cyc, 0.0
        o = O.Oracle(p, xp, yp)
        cf = o.mg_discretize(0.0, o.mg_ngrid())
        ny, nx = cf.shape[1:]
        A, mx, my = _assemble(cf, nx, ny, tuple(p.mudpack_bc))
        rhs = np.random.default_rng(9).standard_normal((ny, nx))
        phi, _ = o.mudpack(0.0, rhs)
        b = -rhs.copy()
        b[ny - 1, :] = 0.0
        res.append(np.abs(A @ phi.ravel() - b.ravel()).max())
    rate = (res[1] / res[0]) ** (1 / 8)
    assert 0.6 < rate < 0.95, rate


def test_discretisation_exact_on_quadratics(oracle_lib):
    """On a uniform rectilinear grid the 9-point operator applied to a
    quadratic in index space equals cxx*2a + cyy*2b + cxy*c + ce*q exactly,
    with cxx = G11 = dy/dz, cyy = G22 = dz/dy, cxy = 0 (src/duct.F90:3917-3931)."""
    O = oracle_lib
    ny, nz, dy, dz = 15, 15, 0.5, 0.25
    xp, yp = np.meshgrid(dz * np.arange(nz + 2), dy * np.arange(ny + 2))
    o, p = _oracle_for_grid(O, ny, nz, xp, yp)
    kx2 = 0.3
    cf = o.mg_discretize(kx2, o.mg_ngrid())
    nyy, nxx = cf.shape[1:]
    i, j = np.meshgrid(np.arange(nxx, dtype=float), np.arange(nyy, dtype=float))
    a, b, c, d, e = 0.3, -0.7, 0.45, 1.1, -0.2
    q = a * i * i + b * j * j + c * i * j + d * i + e * j + 2.0
    off = {0: (1, 0), 1: (1, 1), 2: (0, 1), 3: (-1, 1), 4: (-1, 0), 5: (-1, -1), 6: (0, -1), 7: (1, -1)}
    lhs = cf[8] * q
    for k, (di, dj) in off.items():
        lhs = lhs + cf[k] * np.roll(np.roll(q, -dj, axis=0), -di, axis=1)
    cxx, cyy, J = dy / dz, dz / dy, dy * dz
    expect = cxx * 2 * a + cyy * 2 * b - kx2 * J * q
    inner = (slice(1, nyy - 1), slice(1, nxx - 1))
    assert np.abs(lhs[inner] - expect[inner]).max() < 1e-11


# -- A9: projection invariant ---------------------------------------------------
def test_projection_on_uniform_grid_leaves_residual_divergence(oracle_lib):
    """On a uniform rectilinear grid MUDPACK's operator is the divergence of
    the projection's face gradient, so a tight solve leaves J div u at the
    multigrid residual (src/duct.F90:2713-2800)."""
    O = oracle_lib
    p = base_params(16, 31, 31)
    p.mg_maxcy, p.mg_tolmax = 60, 1e-15
    for s, v in enumerate((2, 2, 2, 1)):
        p.mudpack_bc[s] = v
    xp, yp = tanh_grid(p.ny, p.nz, 4.0, 4.0, 0.0)
    o = O.Oracle(p, xp, yp)
    rng = np.random.default_rng(3)
    shp = p.shape_phys
    u = [rng.standard_normal(shp) for _ in range(3)]
    for a in u:
        a[:, :, p.nx:] = 0.0
    o.set_initial_physical(u[0], u[1], u[2], np.zeros(shp))
    d0 = o.diag().div_ref
    assert d0 < 1e-11


def test_ekman_initial_condition():
    """The seeded initial field is the laminar Ekman spiral plus bounded
    uniform noise; padding entries are zero."""
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    u1, u2, u3, th = ekman_initial(p, yp, cfg.seed, 0.0, 0.0)
    eta = yp[:, :, None]
    assert np.abs(u1[:, :, :p.nx] - (1 - np.exp(-eta) * np.cos(eta))).max() < 1e-15
    assert np.all(u1[:, :, p.nx:] == 0.0)
    n1, _, _, _ = ekman_initial(p, yp, cfg.seed, cfg.noise, 0.0)
    dn = n1[:, :, :p.nx] - u1[:, :, :p.nx]
    assert np.abs(dn).max() <= 0.5 * cfg.noise + 1e-15
    assert np.abs(dn).max() > 0.4 * cfg.noise
    n2, _, _, _ = ekman_initial(p, yp, cfg.seed, cfg.noise, 0.0)
    assert np.array_equal(n1, n2)


# -- proj_mode 1: the consistent projection --------------------------------------
def _tight_oracle(n=1, tol=1e-13):
    cfg = config(n)
    p = cfg.params
    p.proj_mode, p.proj_tol, p.proj_maxit = 1, tol, 60
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    return cfg, p, xp, yp


def test_tight_rows_are_divergence_of_projection_gradient(oracle_lib):
    """On the cells J=2..NY-1, K=2..NZ-1 the operator rows equal J div u of
    the projection's correction of a field phi, computed here directly from
    the reference's face and centre corrections (src/duct.F90:2713-2731) and
    its divergence (:2754-2760), with a skewed grid so every metric term is
    active."""
    O = oracle_lib
    cfg, p, xp, yp = _tight_oracle()
    xp = xp + 0.05 * np.sin(0.3 * yp)
    o = O.Oracle(p, xp, yp)
    kx2 = 0.37
    cf = o.tight_operator(kx2)
    G11, G12, G22, JAC = o.metric(0), o.metric(1), o.metric(2), o.metric(7)
    rng = np.random.default_rng(4)
    phi = rng.standard_normal((p.ny + 2, p.nz + 2))
    NY, NZ = p.ny, p.nz
    js, je, zs, ze = 1, NY, 1, NZ                     # ranges of this BC set
    F2 = np.zeros((NY + 3, NZ + 3))                   # corrections of CU2b(K,J), CU3b(K,J)
    F3 = np.zeros((NY + 3, NZ + 3))
    for J in range(js, je + 1):
        for K in range(zs, ze + 1):
            F2[J, K] = G22[0, J, K] * (phi[J, K] - phi[J - 1, K]) + 0.25 * G12[0, J, K] * (
                phi[J, K + 1] + phi[J - 1, K + 1] - phi[J, K - 1] - phi[J - 1, K - 1])
            F3[J, K] = G11[1, J, K] * (phi[J, K] - phi[J, K - 1]) + 0.25 * G12[1, J, K] * (
                phi[J + 1, K] + phi[J + 1, K - 1] - phi[J - 1, K] - phi[J - 1, K - 1])
    off = {0: (1, 0), 1: (1, 1), 2: (0, 1), 3: (-1, 1), 4: (-1, 0), 5: (-1, -1), 6: (0, -1), 7: (1, -1)}
    worst = 0.0
    for J in range(2, NY):
        for K in range(2, NZ):
            # J div u of the correction: -(i kx J)(i kx phi) + face differences, with the sign
            # the solve uses (L phi = div of the subtracted correction)
            div = (F2[J + 1, K] - F2[J, K]) + (F3[J, K + 1] - F3[J, K]) - kx2 * JAC[J, K] * phi[J, K]
            row = cf[8, J, K] * phi[J, K] + sum(cf[q, J, K] * phi[J + dj, K + di] for q, (di, dj) in off.items())
            worst = max(worst, abs(row - div) / (abs(div) + 1e-300))
    assert worst < 1e-12, worst


@pytest.mark.parametrize("bcs", [(2, 2, 2, 1), (0, 0, 2, 1), (0, 0, 2, 2), (2, 2, 2, 2)])
def test_tight_solve_matches_dense(bcs, oracle_lib):
    """The consistent solve against a sparse direct solve of the same
    operator (its rows from tight_operator, assembled with MUDPACK's virtual
    points for periodic z ends: phi(-1) = phi(nx-2), phi(nx) = phi(1)), for
    the configurations' set, periodic z ends with a specified top, and the
    two sets with no specified side, whose kx = 0 problem is singular: there
    the direct solve replaces the row of the point (K = 0, J = NY+1) by
    phi = 0, the pin the oracle applies, and every other row must hold."""
    O = oracle_lib
    cfg, p, xp, yp = _tight_oracle()
    p.mudpack_bc[:] = bcs
    xp = xp + 0.05 * np.sin(0.3 * yp)                   # every metric term active
    o = O.Oracle(p, xp, yp)
    singular_seen = False
    for kx2 in (0.0, 0.25):
        rhs = np.random.default_rng(3).standard_normal((p.ny + 2, p.nz + 2)) * 1e-3
        phi, it = o.tight_solve(kx2, rhs)
        cf = o.tight_operator(kx2)
        A, _, _ = _assemble(np.concatenate([cf, np.zeros((1,) + cf.shape[1:])]), p.nz + 2, p.ny + 2, bcs)
        b = -rhs.copy()
        if bcs[3] == 1:
            b[-1, :] = 0.0                              # Dirichlet top (MUDPACK_BC(4) = 1)
        if kx2 == 0.0 and 1 not in bcs:
            singular_seen = True
            # constants are in the null space of every row
            assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-9 * np.abs(A).max()
            pin = (p.ny + 1) * (p.nz + 2)
            A = A.tolil()
            A[pin, :] = 0.0
            A[pin, pin] = 1.0
            b[-1, 0] = 0.0
            assert phi[-1, 0] == 0.0
        x = spla.spsolve(A.tocsc(), b.ravel()).reshape(b.shape)
        assert it < 40
        assert rel(phi, x) < 1e-8, (bcs, kx2, rel(phi, x))
    assert singular_seen == (1 not in bcs)


def test_tight_projection_removes_divergence(oracle_lib):
    """proj_mode 1 leaves |J div u| at the solve tolerance on the cells the
    reference's diagnostic covers; the MUDPACK replica does not."""
    O = oracle_lib
    cfg, p, xp, yp = _tight_oracle(1, 1e-11)
    u1, u2, u3, th = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    o = O.Oracle(p, xp, yp)
    o.set_initial_physical(u1, u2, u3, th)
    d = o.diag()
    assert d.div_max < 1e-10 and d.div_ref < 1e-11
    p0 = cfg.params.copy()
    p0.proj_mode = 0
    o0 = O.Oracle(p0, xp, yp)
    o0.set_initial_physical(u1, u2, u3, th)
    assert o0.diag().div_ref > 1e-8


# ---------------------------------------------------------------------------
# LES: dynamic Smagorinsky model of the velocity (src/les_chan.F90 model 2)
# ---------------------------------------------------------------------------
def _les_pair(noise, time_step=3):
    """Oracle objects without and with the subgrid model, both restarted from
    the same projected state of config 1 (a restart, so the model runs from
    the first step and recomputes its coefficient there)."""
    from oracle import oracle as O
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, noise, cfg.th_noise)
    o0 = O.Oracle(p, xp, yp)
    o0.set_initial_physical(*ic)
    st = o0.get_state()
    pl = p.copy()
    pl.les_model = 2
    oL = O.Oracle(pl, xp, yp)
    oL.set_state(*st[:4], st[4], time=0.0, time_step=time_step)
    oD = O.Oracle(p, xp, yp)
    oD.set_state(*st[:4], st[4], time=0.0, time_step=time_step)
    return p, oD, oL


def test_les_vanishes_without_x_fluctuations(oracle_lib):
    """On an x-independent flow (the laminar Ekman spiral) the test-filtered
    products equal the products of the filtered velocities, so the dynamic
    coefficient is zero up to rounding and a substep with the model is the
    plain substep (fields agree to 1e-12 relative L2)."""
    p, oD, oL = _les_pair(noise=0.0)
    oD.rk_duct()
    oL.rk_duct()
    nut, cdyn = oL.get_les()
    assert np.abs(cdyn).max() < 1e-10
    assert np.abs(nut).max() < 1e-12
    for a, b in zip(oL.get_state(), oD.get_state()):
        n = np.linalg.norm(b)
        assert np.linalg.norm(a - b) <= 1e-12 * (n if n > 0 else 1.0)


def test_les_coefficient_and_viscosity_signs(oracle_lib):
    """With resolved x fluctuations (5% noise) the clipped coefficient is
    nonnegative, positive somewhere, and zero outside the interior index
    ranges; NU_T = C_DYN Delta^2 |S| is nonnegative and zero outside the
    interior (src/les_chan.F90:798-831)."""
    p, oD, oL = _les_pair(noise=0.05)
    oL.rk_duct()
    nut, cdyn = oL.get_les()
    assert (cdyn >= 0.0).all() and (nut >= 0.0).all()
    assert cdyn.max() > 0.0 and nut.max() > 0.0
    # config 1: interior rows J = 1..NY, K = 1..NZ (ranges of its BC types)
    assert np.all(cdyn[0, :] == 0.0) and np.all(cdyn[-1, :] == 0.0)
    assert np.all(cdyn[:, 0] == 0.0) and np.all(cdyn[:, -1] == 0.0)
    assert np.all(nut[0] == 0.0) and np.all(nut[-1] == 0.0)
    # the model changes the step
    a, b = oL.get_state()[0], None
    oD.rk_duct()
    b = oD.get_state()[0]
    assert np.linalg.norm(a - b) > 0.0


def _les_th_pair(noise, th_noise, shift=0.0, time_step=3):
    """Oracle objects without and with the subgrid models on a small
    stratified grid (config 2's physics, 32 x 33 x 33 points), both restarted
    from the same projected state; `shift` adds a constant to the scalar."""
    from oracle import oracle as O
    from ekman_bl_amd.params import mg_factors
    cfg = config(2)
    p = cfg.params
    p.nx, p.ny, p.nz = 32, 31, 31
    p.mg_ixp, p.mg_iex = mg_factors(p.nz + 2)
    p.mg_jyq, p.mg_jey = mg_factors(p.ny + 2)
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, noise, th_noise)
    o0 = O.Oracle(p, xp, yp)
    o0.set_initial_physical(*ic)
    st = o0.get_state()
    cth = st[4].copy()
    cth[..., 0] += shift
    pl = p.copy()
    pl.les_model = 2
    oL = O.Oracle(pl, xp, yp)
    oL.set_state(*st[:4], cth, time=0.0, time_step=time_step)
    oD = O.Oracle(p, xp, yp)
    oD.set_state(*st[:4], cth, time=0.0, time_step=time_step)
    return p, oD, oL


def test_les_th_vanishes_without_x_fluctuations(oracle_lib):
    """The scalar model (src/les_chan_th.F90, model 2) on an x-independent
    stratified flow: the filtered products equal the products of the filtered
    fields, so C_DYN_TH and KAPPA_T vanish up to rounding and a substep with
    both models is the plain substep for every field, scalar included."""
    p, oD, oL = _les_th_pair(0.0, 0.0)
    oD.rk_duct()
    oL.rk_duct()
    kap, cdt = oL.get_les_th()
    assert np.abs(cdt).max() < 1e-10 and np.abs(kap).max() < 1e-12
    for a, b in zip(oL.get_state(), oD.get_state()):
        n = np.linalg.norm(b)
        assert np.linalg.norm(a - b) <= 1e-12 * (n if n > 0 else 1.0)


def test_les_th_coefficient_and_diffusivity_signs(oracle_lib):
    """With x fluctuations of velocity and scalar the clipped C_DYN_TH is
    nonnegative, positive somewhere and zero outside the interior rows;
    KAPPA_T = C_DYN_TH Delta^2 |S| is nonnegative and zero outside the
    scalar's interior (src/les_chan_th.F90:300-334); the model changes the
    scalar."""
    p, oD, oL = _les_th_pair(0.05, 0.01)
    oL.rk_duct()
    kap, cdt = oL.get_les_th()
    assert (cdt >= 0.0).all() and (kap >= 0.0).all()
    assert cdt.max() > 0.0 and kap.max() > 0.0
    assert np.all(cdt[0, :] == 0.0) and np.all(cdt[-1, :] == 0.0)
    assert np.all(cdt[:, 0] == 0.0) and np.all(cdt[:, -1] == 0.0)
    assert np.all(kap[0] == 0.0) and np.all(kap[-1] == 0.0)
    assert np.all(kap[:, 0] == 0.0) and np.all(kap[:, -1] == 0.0)
    oD.rk_duct()
    a, b = oL.get_state()[4], oD.get_state()[4]
    assert np.linalg.norm(a - b) > 1e-9 * np.linalg.norm(b)


def test_les_th_coefficient_invariant_under_scalar_shift(oracle_lib):
    """A constant added to the scalar leaves its gradient unchanged (the
    difference operators annihilate constants on the rectilinear grid) and
    cancels in the resolved flux filter(u_i th) - filter(u_i) filter(th), so
    C_DYN_TH and KAPPA_T of the first substep do not change (to rounding).
    A filtered velocity paired with the wrong product would break this."""
    _, _, o1 = _les_th_pair(0.05, 0.01)
    _, _, o2 = _les_th_pair(0.05, 0.01, shift=3.0)
    o1.rk_duct()
    o2.rk_duct()
    k1, c1 = o1.get_les_th()
    k2, c2 = o2.get_les_th()
    assert c1.max() > 0.0
    assert np.abs(c1 - c2).max() <= 1e-8 * c1.max()
    assert np.abs(k1 - k2).max() <= 1e-8 * k1.max()


def test_flow_rate_controller_closed_form(oracle_lib):
    """F_TYPE 0 (src/duct.F90:414-436, constants src/diablo.F90:359-361): with
    a uniform mean W = c on a grid whose y lines do not depend on z, the flow
    rate integral is c, so one substep moves the pressure gradient by
    -(WEIGHT/TEMP4)(Qn + Q0 - 2c), WEIGHT the single-precision 0.1 the
    reference assigns, TEMP4 = H_BAR(1) = 8/15 DELTA_T, Qn = Q0 = WBULK0."""
    from oracle import oracle as O
    cfg = config(1)
    p = cfg.params
    p.f_type, p.wbulk0, p.px0, p.variable_dt = 0, 0.3, 0.01, 0
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    o = O.Oracle(p, xp, yp)
    z = np.zeros(p.shape_fourier, dtype=np.complex128)
    cu3 = z.copy()
    c = 0.5
    cu3[..., 0] = c
    o.set_state(z, z, cu3, z, z, time=0.0, time_step=0)
    assert o.forcing() == (0.01, 0.3)
    o.rk_duct()
    px0, qn = o.forcing()
    t4 = p.delta_t * (8.0 / 15.0)
    w = float(np.float32(0.1))
    assert abs(qn - c) <= 1e-14
    assert abs(px0 - (0.01 - (w / t4) * (0.3 + 0.3 - 2 * c))) <= 1e-13
    # the gradient acts on the mean mode: a second substep keeps integrating
    o.rk_duct()
    px1, qn1 = o.forcing()
    assert px1 != px0 and abs(qn1 - c) < 0.5


def test_reference_ic7_form():
    """IC_TYPE 7 (src/duct.F90:3026-3137, scalar :2854-2864): U3 = WBULK0 in
    the mean mode everywhere, U1 = U2 = TH = 0 there; the kicks only touch the
    real parts of modes 1..NX/2 at the interior points, within
    +-KICK/2 * exp(-6 J^2/NY^2), and fill that range; the physical fields
    transform back to the same modes (forward scaled 1/NX, modes <= NX/3)."""
    from ekman_bl_amd import ic
    p = config(1).params
    p.wbulk0 = 0.7
    kick = 0.1
    cu1, cu2, cu3, cp, cth = ic.reference_ic7_fourier(p, kick, 3)
    assert np.all(cu3[..., 0] == 0.7) and np.all(cu1[..., 0] == 0) and np.all(cu2[..., 0] == 0)
    assert not cp.any() and not cth.any()
    js, je, zs, ze = ic.index_ranges(p)
    assert (js, je, zs, ze) == (1, p.ny, 1, p.nz)   # config 1: W no-slip bottom, Neumann z ends
    for c in (cu1, cu2, cu3):
        kicks = c[..., 1:]
        assert not kicks.imag.any()
        inner = np.zeros(kicks.shape, bool)
        inner[js:je + 1, zs:ze + 1, :] = True
        assert not kicks[~inner].any()
        j = np.arange(p.ny + 2)[:, None, None]
        bound = 0.5 * kick * np.exp(-6.0 * j ** 2 / p.ny ** 2)
        assert np.all(np.abs(kicks.real) <= bound + 1e-15)
        # uniform draws: the damped amplitude fills its range near the wall
        assert np.abs(kicks.real[1]).max() > 0.9 * bound[1, 0, 0]
    u1, u2, u3, th = ic.reference_ic7(p, kick, 3)
    back = np.fft.rfft(u3[..., :p.nx], axis=-1) / p.nx
    m = p.nx // 3 + 1
    assert np.abs(back[..., :m] - cu3[..., :m]).max() < 1e-14
    assert np.abs(back[..., m:]).max() < 1e-14
    assert not th.any()


def test_x_statistics_single_mode_closed_form(oracle_lib):
    """SAVE_STATS_CURVI's x moments (src/flow_statistics.F90:163-215, 420-495)
    as the oracle forms them (direct sums over the NX physical points): with
    mode 1 of U1 equal to a and of U2 equal to b (u = 2a cos x, v = 2|b|
    cos(x + arg b)) and means c, d, the sums are urms = 2 NX a^2, vrms =
    2 NX |b|^2, uv = 2 Re(a conj b), wrms = uw = 0, and the means are c, d."""
    from oracle import oracle as O
    cfg = config(1)
    p = cfg.params
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    o = O.Oracle(p, xp, yp)
    z = np.zeros(p.shape_fourier, dtype=np.complex128)
    a, b, c, d = 0.3, 0.2 - 0.1j, 1.5, -0.25
    cu1, cu2 = z.copy(), z.copy()
    cu1[..., 1], cu2[..., 1] = a, b
    cu1[..., 0], cu2[..., 0] = c, d
    o.set_state(cu1, cu2, z, z, z, time=0.0, time_step=0)
    st = o.x_stats()
    nx = p.nx
    assert np.allclose(st["u1"], c, rtol=0, atol=1e-15) and np.allclose(st["u2"], d, rtol=0, atol=1e-15)
    assert np.allclose(st["urms"], 2 * nx * a * a, rtol=1e-13, atol=0)
    assert np.allclose(st["vrms"], 2 * nx * abs(b) ** 2, rtol=1e-13, atol=0)
    assert np.allclose(st["uv"], 2 * (a * np.conj(b)).real, rtol=1e-13, atol=0)
    assert np.abs(st["wrms"]).max() == 0 and np.abs(st["uw"]).max() == 0


# ---------------------------------------------------------------------------
# sponge_th (src/duct.F90:3196-3362) and its profile (src/diablo.F90:1533-1556)
# ---------------------------------------------------------------------------
def _sponge_small(sponge_type=2, ny_s_t=None, ri_tau=0.0):
    """A 16 x 33 x 17 stratified-config case with a vertical sponge on the top
    rows, the plain projection, and the z rows of config 2 (not type 6, so the
    sponge's type-6 branch stays off); ri_tau = 0 keeps the scalar passive."""
    from ekman_bl_amd.params import mg_factors
    cfg = config(2)
    p = cfg.params
    p.nx, p.ny, p.nz = 16, 31, 15
    p.mg_ixp, p.mg_iex = mg_factors(p.nz + 2)
    p.mg_jyq, p.mg_jey = mg_factors(p.ny + 2)
    p.proj_mode = 0
    p.ri_tau = ri_tau
    p.sponge_type = sponge_type
    p.ny_s_t = p.ny - 7 if ny_s_t is None else ny_s_t
    p.nz_s_r = p.nz + 2
    p.save_stats_int = 1 << 30
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    return cfg, p, xp, yp


@pytest.mark.parametrize("sponge_type", [1, 2])
def test_sponge_profile_closed_forms(sponge_type, oracle_lib):
    """SPONGE_SIGMA is 0 below and at NY_S_T and 1 at NY+1 for both types;
    type 1 is the square of the height above ypoint(NZ+1, NY_S_T) over the
    span (the span passes through REAL(), single precision), so on a grid
    whose y varies with K only the column K = NZ+1 reproduces it; type 2 has
    sigma + 1/29 growing by the constant factor 30^(1/(NY+1-NY_S_T)) per row.
    The library's host profile equals the oracle's bit for bit."""
    from ekman_bl_amd import solver as S
    O = oracle_lib
    cfg, p, xp, yp = _sponge_small(sponge_type)
    p.n_th = 1
    jf = np.arange(p.ny + 2)[:, None] / (p.ny + 1)
    yp = yp + 0.2 * p.ly * (xp / p.lz) * jf ** 2  # y depends on K, nonlinearly in J
    o = O.Oracle(p, xp, yp)
    sig = o.sponge_sigma()
    nys, ny = p.ny_s_t, p.ny
    assert np.all(sig[:nys + 1] == 0.0)
    assert np.all(np.diff(sig[nys:]) > 0)
    if sponge_type == 1:
        assert abs(sig[ny + 1] - 1.0) < 5e-7
        span = np.float32(yp[ny + 1, p.nz + 1] - yp[nys, p.nz + 1])
        h = (yp[nys:, p.nz + 1] - yp[nys, p.nz + 1]) / float(span)
        np.testing.assert_allclose(sig[nys:], h * h, rtol=1e-14, atol=0)
        h0 = (yp[nys:, 0] - yp[nys, 0]) / (yp[ny + 1, 0] - yp[nys, 0])
        assert np.max(np.abs(sig[nys:] - h0 * h0)) > 1e-3
    else:
        assert abs(sig[ny + 1] - 1.0) < 1e-14
        ratio = (sig[nys + 1:] + 1 / 29) / (sig[nys:-1] + 1 / 29)
        np.testing.assert_allclose(ratio, 30.0 ** (1.0 / (ny + 1 - nys)), rtol=1e-13)
    host = S.host_sponge_sigma(sponge_type, p.nz, p.ny, nys, yp)
    assert np.array_equal(host, sig)


def _oracle_run(O, p, xp, yp, state=None, ic=None, substeps=1):
    o = O.Oracle(p, xp, yp)
    if state is not None:
        o.set_state(*state)
    else:
        o.set_initial_physical(*ic)
    for _ in range(substeps):
        assert o.rk_duct() == 0
    return o.get_state()


def test_sponge_without_rows_is_the_identity(oracle_lib):
    """With NY_S_T above NY+1 every SPONGE_SIGMA is zero and, without the
    type-6 branch, sponge_th changes nothing: the run equals the run without
    a sponge bit for bit (the mean-U3 average then covers J >= 2 NY/3)."""
    O = oracle_lib
    cfg, p, xp, yp = _sponge_small(2, ny_s_t=1 << 20)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    a = _oracle_run(O, p, xp, yp, ic=ic, substeps=3)
    p.sponge_type = 0
    b = _oracle_run(O, p, xp, yp, ic=ic, substeps=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_sponge_leaves_a_state_at_its_targets_alone(oracle_lib):
    """A state whose fluctuations vanish and whose mean velocities equal the
    sponge's targets (CU1 -> V_BC_YMAX_C1, CU2 -> U_BC_YMAX_C1, CU3 -> WBULK0,
    src/duct.F90:3279-3281) is a fixed point of the damping: one substep
    equals the substep without a sponge bit for bit.  Swapping two targets
    breaks it (checked by moving U_BC_YMAX_C1)."""
    from ekman_bl_amd.params import YMAX
    O = oracle_lib
    cfg, p, xp, yp = _sponge_small(1)
    p.u_bc_c1[1][YMAX] = 0.3      # V_BC_YMAX_C1: target of CU1
    p.u_bc_c1[0][YMAX] = -0.2     # U_BC_YMAX_C1: target of CU2
    p.wbulk0 = 0.7
    shape = (p.ny + 2, p.nz + 2, p.nx // 2 + 1)
    fl = [np.zeros(shape, complex) for _ in range(5)]
    fl[0][..., 0], fl[1][..., 0], fl[2][..., 0] = 0.3, -0.2, 0.7
    fl[4][..., 0] = 0.01 * np.arange(p.ny + 2)[:, None]
    a = _oracle_run(O, p, xp, yp, state=fl)
    p.sponge_type = 0
    b = _oracle_run(O, p, xp, yp, state=fl)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    p.sponge_type = 1
    p.u_bc_c1[0][YMAX] = 0.3
    c = _oracle_run(O, p, xp, yp, state=fl)
    assert not np.array_equal(c[1], b[1])


def test_sponge_damps_the_top_fluctuations(oracle_lib):
    """Rayleigh damping with Sponge_Amp = 0.1/DELTA_T: after two steps the
    fluctuation energy (modes >= 1) on the sponge rows is lower than without
    the sponge, and the rows below NY_S_T - 4 are nearly untouched."""
    O = oracle_lib
    cfg, p, xp, yp = _sponge_small(2)
    ic = ekman_initial(p, yp, cfg.seed, 0.2, cfg.th_noise)
    a = _oracle_run(O, p, xp, yp, ic=ic, substeps=6)
    p.sponge_type = 0
    b = _oracle_run(O, p, xp, yp, ic=ic, substeps=6)
    top = slice(p.ny_s_t + 2, p.ny + 1)
    low = slice(1, p.ny_s_t - 4)
    for f in range(3):
        ea = np.sum(np.abs(a[f][top, :, 1:]) ** 2)
        eb = np.sum(np.abs(b[f][top, :, 1:]) ** 2)
        print(f"[sponge] field {f}: top fluctuation energy {ea:.3e} with, {eb:.3e} without", flush=True)
        assert ea < 0.97 * eb
    ra = np.sum(np.abs(a[0][low] - b[0][low]) ** 2) / np.sum(np.abs(b[0][low]) ** 2)
    assert ra < 1e-3


# ---------------------------------------------------------------------------
# SAVE_STATS_CURVI mean gradients (src/flow_statistics.F90:164-175, 228-275, 455-553)
# ---------------------------------------------------------------------------
def test_mean_gradients_closed_forms(oracle_lib):
    """On a sheared affine grid (z = dz K + s y, y = dy J) the cell-averaged
    contravariant metrics times central index differences give the exact
    derivatives of linear fields: with U1 = a y + b z, U3 = c y + e z and
    theta = g y + h z the planes are dudy = a, dudz = b, dwdy = c, dwdz = e,
    dthdy = g, dthdz = h (to the metrics' 1e-8 truncation); a swapped metric
    pair (12 for 21) or a sign fails on the shear.  With CONT_STRAT,
    Rig = -RI_TAU (g - 1)/(a^2 + c^2); TKE = U1^2 + U2^2 + U3^2; the scalar
    rows extrapolate linearly (still g); THBULK is the mean over the cells of
    theta at the cell corner (K-1/2, J-1/2)."""
    O = oracle_lib
    ny, nz, dy, dz, sh = 15, 15, 0.5, 0.25, 0.3
    J = np.arange(ny + 2)[:, None] * np.ones((1, nz + 2))
    K = np.ones((ny + 2, 1)) * np.arange(nz + 2)[None, :]
    yp = dy * J
    xp = dz * K + sh * yp
    p = base_params(16, ny, nz, n_th=1)
    p.cont_strat, p.ri_tau = 1, 0.04
    o = O.Oracle(p, xp, yp)
    a, b, c, e, g, h = 0.7, -0.4, 0.3, 0.9, -0.2, 0.5
    F = [np.zeros((ny + 2, nz + 2, 9), complex) for _ in range(5)]
    F[0][..., 0] = a * yp + b * xp
    F[1][..., 0] = 0.1
    F[2][..., 0] = c * yp + e * xp
    F[4][..., 0] = g * yp + h * xp
    o.set_state(*F)
    m, thb = o.mean_grads()
    inner = (slice(2, ny), slice(2, nz))
    for name, v in (("dudy", a), ("dudz", b), ("dwdy", c), ("dwdz", e), ("dthdy", g), ("dthdz", h)):
        np.testing.assert_allclose(m[name][inner], v, rtol=0, atol=1e-6, err_msg=name)
    np.testing.assert_allclose(m["rig"][inner], -p.ri_tau * (g - 1.0) / (a * a + c * c), rtol=1e-5)
    np.testing.assert_allclose(m["tke"], F[0][..., 0].real ** 2 + 0.01 + F[2][..., 0].real ** 2, rtol=1e-14)
    # fills: velocity rows 0 and NY+1 copy rows 1 and NY; scalar rows extrapolate
    assert np.array_equal(m["dudy"][0, 1:nz + 1], m["dudy"][1, 1:nz + 1])
    assert np.array_equal(m["dwdz"][ny + 1, 2:nz + 1], m["dwdz"][ny, 2:nz + 1])  # column 1 is rewritten after
    np.testing.assert_allclose(m["dthdy"][0, 2:nz], g, atol=1e-6)
    np.testing.assert_allclose(m["dthdz"][ny + 1, 2:nz], h, atol=1e-6)
    yc = dy * (np.arange(1, ny + 1) - 0.5)[:, None]
    zc = dz * (np.arange(1, nz + 1) - 0.5)[None, :] + sh * yc
    np.testing.assert_allclose(thb, np.mean(g * yc + h * zc), rtol=1e-12)


@pytest.mark.parametrize("sponge_type", [1, 2])
def test_sponge_out_profile_closed_forms(sponge_type, oracle_lib):
    """SPONGE_SIGMA_OUT (src/diablo.F90:1489-1523): zero between NZ_S_L and
    NZ_S_R, 0 at both and 1 at NZ+1; type 1 the squared distance along
    xpoint(K, NY+1) over the right span (single precision), read on the top
    row of a grid whose z varies with J, so on a uniform top row the two sides
    mirror each other; type 2 a constant growth factor 25^(1/(NZ+1-NZ_S_R))
    of sigma + 1/24 on both sides.  The library's host profile equals the
    oracle's bit for bit; NZ_S_R >= NZ+1 turns it off."""
    from ekman_bl_amd import solver as S
    O = oracle_lib
    cfg, p, xp, yp = _sponge_small(sponge_type)
    p.n_th = 1
    p.nz_s_r = p.nz - 4
    p.nz_s_l = p.nz + 1 - p.nz_s_r
    jf = np.arange(p.ny + 2)[:, None] / (p.ny + 1)
    xp = xp + 0.3 * np.sin(2.0 * np.pi * jf) * (xp / p.lz) ** 2  # nonlinear in z except at J = 0, NY+1
    o = O.Oracle(p, xp, yp)
    so = o.sponge_sigma_out()
    zr, zl, nz = p.nz_s_r, p.nz_s_l, p.nz
    assert np.all(so[zl:zr + 1] == 0.0)
    assert np.all(np.diff(so[zr:]) > 0) and np.all(np.diff(so[:zl + 1]) < 0)
    if sponge_type == 1:
        assert abs(so[nz + 1] - 1.0) < 5e-7
        xt = xp[p.ny + 1]
        span = float(np.float32(xt[nz + 1] - xt[zr]))
        np.testing.assert_allclose(so[zr:], ((xt[zr:] - xt[zr]) / span) ** 2, rtol=1e-14, atol=0)
        np.testing.assert_allclose(so[:zl + 1][::-1], so[zr:], rtol=1e-12, atol=1e-15)
        xm = xp[p.ny // 2]
        assert np.max(np.abs(so[zr:] - ((xm[zr:] - xm[zr]) / (xm[nz + 1] - xm[zr])) ** 2)) > 1e-4
    else:
        assert abs(so[nz + 1] - 1.0) < 1e-14
        ratio = (so[zr + 1:] + 1 / 24) / (so[zr:-1] + 1 / 24)
        np.testing.assert_allclose(ratio, 25.0 ** (1.0 / (nz + 1 - zr)), rtol=1e-13)
        np.testing.assert_allclose(so[:zl + 1][::-1], so[zr:], rtol=1e-15)
    host = S.host_sponge_sigma_out(sponge_type, p.nz, p.ny, zr, zl, xp)
    assert np.array_equal(host, so)
    p.nz_s_r = p.nz + 2
    assert np.all(O.Oracle(p, xp, yp).sponge_sigma_out() == 0.0)


def test_x_moments_single_mode_closed_forms(oracle_lib):
    """high_oder_momentums_curvi (src/flow_statistics.F90:1026-1095) on a state
    whose fluctuations are one x mode m: u' = a cos, v' = b cos, w' = c cos
    (CU = amplitude/2 at mode m), means U, V, W on mode 0.  Averages over the
    NX points: cos^2 -> 1/2, cos^3 -> 0, cos^4 -> 3/8; so UU = a^2/2,
    UUUU = 3 a^4/8, UUU = UUV = 0, UUUV = 3 a^3 b/8, WWWV = 3 c^3 b/8.  The
    means must be removed first (a mean left in fails UUU)."""
    O = oracle_lib
    p = base_params(16, 7, 7)
    ny, nz = p.ny, p.nz
    y = 0.5 * np.arange(ny + 2)
    z = 0.25 * np.arange(nz + 2)
    xp, yp = np.meshgrid(z, y)
    o = O.Oracle(p, xp, yp)
    a, b, c, m = 0.6, -0.3, 0.9, 3
    F = [np.zeros((ny + 2, nz + 2, 9), complex) for _ in range(5)]
    for f, (amp, mean) in enumerate(((a, 1.5), (b, -0.2), (c, 0.7))):
        F[f][..., 0] = mean
        F[f][..., m] = amp / 2
    o.set_state(*F)
    x = o.x_moments()
    exp = {"uu": a * a / 2, "vv": b * b / 2, "ww": c * c / 2, "uuu": 0.0, "vvv": 0.0, "www": 0.0,
           "uuuu": 3 * a ** 4 / 8, "vvvv": 3 * b ** 4 / 8, "wwww": 3 * c ** 4 / 8, "uuv": 0.0, "wwv": 0.0,
           "uuuv": 3 * a ** 3 * b / 8, "wwwv": 3 * c ** 3 * b / 8}
    for name, v in exp.items():
        np.testing.assert_allclose(x[name], v, rtol=0, atol=1e-14, err_msg=name)
