"""The mean streamwise vorticity of the oracle and the plane, 3D and bulk
writers of SAVE_STATS_CURVI (ekman_bl_amd/planes.py) on the CPU, against
closed forms: linear fields on affine grids, where the derivatives the
reference forms are exact or have a known error."""
import os

import numpy as np
import pytest

from ekman_bl_amd import planes as PL
from ekman_bl_amd.params import base_params
from ekman_bl_amd.restart import read_record


def test_omega_x_closed_form(oracle_lib):
    """On a sheared affine grid (z = dz K + s y, y = dy J) the cell-averaged
    contravariant metrics times central index differences give the exact
    derivatives of linear fields: U2 = a y + b z, U3 = c y + e z give
    omega_x = dW/dy - dV/dz = c - b; a swapped metric pair (12 for 21) or a
    sign fails on the shear.  Rows 0 and NY+1 copy rows 1 and NY; columns 0
    and NZ+1 are zero (src/flow_statistics.F90:301-327)."""
    O = oracle_lib
    ny, nz, dy, dz, sh = 15, 15, 0.5, 0.25, 0.3
    J = np.arange(ny + 2)[:, None] * np.ones((1, nz + 2))
    K = np.ones((ny + 2, 1)) * np.arange(nz + 2)[None, :]
    yp = dy * J
    xp = dz * K + sh * yp
    p = base_params(16, ny, nz, n_th=0)
    o = O.Oracle(p, xp, yp)
    a, b, c, e = 0.7, -0.4, 0.3, 0.9
    F = [np.zeros((ny + 2, nz + 2, 9), complex) for _ in range(5)]
    F[0][..., 0] = 0.2 * yp
    F[1][..., 0] = a * yp + b * xp
    F[2][..., 0] = c * yp + e * xp
    o.set_state(*F[:4], None)
    w = o.omega_x()
    np.testing.assert_allclose(w[2:ny, 2:nz], c - b, rtol=0, atol=1e-6)
    assert np.array_equal(w[0, 1:nz + 1], w[1, 1:nz + 1])
    assert np.array_equal(w[ny + 1, 1:nz + 1], w[ny, 1:nz + 1])
    assert not np.any(w[:, 0]) and not np.any(w[:, nz + 1])


def _records(path):
    with open(path, "rb") as f:
        out = []
        while True:
            try:
                out.append(read_record(f))
            except (EOFError, ValueError):
                break
        return out


def _arr(b, *shape):
    return np.frombuffer(b, dtype="<f8").reshape(shape)


def test_xy_plane_writer(tmp_path):
    """plane_XY_binary_combined: header (TIME, NX, NY, K, NXP), the grids,
    then U1, U2, U3, P, TH and omg_z over (0:NX, 0:NY+1) with the corner
    column NX a copy of column 0.  omg_z = dV/dx - dU/dy by the writer's
    difference, which drops the centre term of the non-uniform three-point
    formula: for U1 = a y it gives a (1 + y (h2 - h1)/(h1 h2)) with h1, h2
    the spacings above and below, exact only where the grid is uniform; the x
    difference at I = NX-1 reaches the corner column (the periodic copy of
    column 0)."""
    nx, ny, nz, k = 12, 9, 7, 3
    y = np.cumsum(np.r_[0.0, 0.1 * 1.2 ** np.arange(ny + 1)])        # stretched, NY+2 points
    yp = np.broadcast_to(y[:, None], (ny + 2, nz + 2)).copy()
    dx = 0.5
    x = np.arange(nx) * dx
    a, bb = 0.8, -0.3
    planes = np.zeros((5, ny + 2, nx))
    planes[0] = a * y[:, None] + 0.0 * x[None, :]
    planes[1] = bb * x[None, :] + 0.0 * y[:, None]
    planes[2] = 1.0
    planes[3] = np.sin(x)[None, :] + y[:, None]
    planes[4] = 2.0
    path = str(tmp_path / "xy.pln")
    PL.write_xy_combined(path, 1.5, k, yp, dx, planes, PL.nxp(nx, 4))
    r = _records(path)
    assert len(r) == 3
    t, = np.frombuffer(r[0][:8], "<f8")
    hdr = np.frombuffer(r[0][8:], "<i4")
    assert t == 1.5 and list(hdr) == [nx, ny, k, PL.nxp(nx, 4)]
    g = _arr(r[1], 2, ny + 2, nx + 1)
    assert np.array_equal(g[0][:, 0], y) and np.array_equal(g[1][0], np.arange(nx + 1) * dx)
    d = _arr(r[2], 6, ny + 2, nx + 1)
    assert np.array_equal(d[:5, :, :nx], planes) and np.array_equal(d[:5, :, nx], planes[:, :, 0])
    h1 = y[2:ny + 2] - y[1:ny + 1]
    h2 = y[1:ny + 1] - y[0:ny]
    dudy = a * (1.0 + y[1:ny + 1] * (h2 - h1) / (h1 * h2))
    want = bb - dudy[:, None]
    np.testing.assert_allclose(d[5, 1:ny + 1, 1:nx - 1], np.broadcast_to(want, (ny, nx - 2)), rtol=1e-12, atol=1e-12)
    # at I = NX-1 the x difference reaches the corner column, the periodic copy of column 0
    np.testing.assert_allclose(d[5, 1:ny + 1, nx - 1], (0.0 - bb * (nx - 2) * dx) / (2 * dx) - dudy, rtol=1e-12)
    assert not np.any(d[5, 0]) and not np.any(d[5, :, 0]) and not np.any(d[5, :, nx])


def test_xz_plane_writer(tmp_path):
    """plane_XZ_binary_combined over (0:NX-1, 1:NZ): omg_y = dU/dz - dW/dx
    (uniform spacings: exact for linear fields) with the reference's
    periodic-looking fills (column 0 from NX-2, column NX-1 from 1, row NZ
    from 2, row 1 from NZ-1); the fluctuations subtract each x point's z mean,
    except U1's, which subtracts the z mean of x point 0 everywhere."""
    nx, ny, nz, j = 10, 5, 8, 2
    dz, dx = 0.3, 0.5
    xp = np.broadcast_to((dz * np.arange(nz + 2))[None, :], (ny + 2, nz + 2)).copy()
    x = np.arange(nx) * dx
    z = dz * np.arange(nz + 2)
    a, c, dd = 0.6, -0.2, 0.4
    planes = np.zeros((5, nz + 2, nx))
    planes[0] = a * z[:, None] + dd * x[None, :]
    planes[1] = np.cos(z)[:, None] * (1.0 + x[None, :])
    planes[2] = c * x[None, :] + 0.0 * z[:, None]
    planes[3] = 1.0
    planes[4] = z[:, None] ** 2 + 0.0 * x[None, :]
    path = str(tmp_path / "xz.pln")
    PL.write_xz_combined(path, 2.0, j, xp, dx, planes, 7)
    r = _records(path)
    assert len(r) == 4
    assert list(np.frombuffer(r[0][8:], "<i4")) == [nx, nz, j, 7]
    g = _arr(r[1], 2, nz, nx)
    assert np.array_equal(g[0][:, 0], z[1:nz + 1]) and np.array_equal(g[1][0], x)
    d = _arr(r[2], 5, nz, nx)
    assert np.array_equal(d, planes[:, 1:nz + 1])
    f = _arr(r[3], 7, nz, nx)
    om = f[5]
    np.testing.assert_allclose(om[1:nz - 1, 1:nx - 1], a - c, rtol=1e-12)
    assert np.array_equal(om[:, 0], om[:, nx - 2]) and np.array_equal(om[:, nx - 1], om[:, 1])
    assert np.array_equal(om[nz - 1], om[1]) and np.array_equal(om[0], om[nz - 2])
    zm = z[1:nz + 1]
    np.testing.assert_allclose(f[0], a * (zm - zm.mean())[:, None] + dd * x[None, :], rtol=0, atol=1e-12)
    np.testing.assert_allclose(f[1], planes[1, 1:nz + 1] - planes[1, 1:nz + 1].mean(axis=0), rtol=0, atol=1e-12)
    np.testing.assert_allclose(f[4], planes[4, 1:nz + 1] - np.mean(zm ** 2), rtol=0, atol=1e-12)
    # omg_y_p = dU1'/dz - dW'/dx: U1' keeps its z slope a; W = c x is its own z mean, so W' = 0
    np.testing.assert_allclose(f[6][1:nz - 1, 1:nx - 1], a, rtol=1e-12)


def test_zy_and_3d_writers(tmp_path):
    """plane_ZY_binary: one record TIME, xpoint, ypoint, U3, U2, U1, P, TH;
    plot_3D: (xpoint, ypoint, dx, U3), U2, U1, TH + TH_BAR, P over x points
    0..NX-1."""
    nx, ny, nz = 6, 4, 5
    rng = np.random.default_rng(3)
    xp, yp = rng.random((ny + 2, nz + 2)), rng.random((ny + 2, nz + 2))
    planes = rng.random((5, ny + 2, nz + 2))
    path = str(tmp_path / "zy.pln")
    PL.write_zy(path, 0.25, xp, yp, planes)
    r = _records(path)
    assert len(r) == 1
    v = np.frombuffer(r[0], "<f8")
    assert v[0] == 0.25
    body = v[1:].reshape(7, ny + 2, nz + 2)
    for got, want in zip(body, (xp, yp, planes[2], planes[1], planes[0], planes[3], planes[4])):
        assert np.array_equal(got, want)
    fields = [rng.random((ny + 2, nz + 2, nx + 2)) for _ in range(5)]
    tb = rng.random((ny + 2, nz + 2))
    path = str(tmp_path / "d3.pln_001")
    PL.write_3d(path, nx, xp, yp, 0.125, fields, tb)
    r = _records(path)
    assert len(r) == 5
    n = (ny + 2) * (nz + 2)
    v0 = np.frombuffer(r[0], "<f8")
    assert np.array_equal(v0[:n].reshape(ny + 2, nz + 2), xp) and v0[2 * n] == 0.125
    assert np.array_equal(v0[2 * n + 1:].reshape(ny + 2, nz + 2, nx), fields[2][:, :, :nx])
    assert np.array_equal(_arr(r[1], ny + 2, nz + 2, nx), fields[1][:, :, :nx])
    assert np.array_equal(_arr(r[2], ny + 2, nz + 2, nx), fields[0][:, :, :nx])
    assert np.array_equal(_arr(r[3], ny + 2, nz + 2, nx), fields[4][:, :, :nx] + tb[:, :, None])
    assert np.array_equal(_arr(r[4], ny + 2, nz + 2, nx), fields[3][:, :, :nx])
    PL.write_3d(path, nx, xp, yp, 0.125, fields[:4] + [None])
    assert len(_records(path)[3]) == 0


def test_bulk_and_time_bulk_line():
    """UBULK/VBULK are the means over the cells of the mode-0 values at the
    cell corners (K-1/2, J-1/2): a constant gives itself, U = K gives
    NZ/2; FORMAT(6f15.7) writes fixed 15-wide fields and asterisks when a
    value does not fit."""
    ny, nz = 7, 9
    kk = np.broadcast_to(np.arange(nz + 2, dtype=float)[None, :], (ny + 2, nz + 2))
    ub, vb = PL.bulk(np.full((ny + 2, nz + 2), 0.75), kk)
    assert ub == pytest.approx(0.75, rel=1e-15) and vb == pytest.approx(nz / 2, rel=1e-14)
    line = PL.time_bulk_line(12.5, 0.1, -0.25, 3.0, 1e9, -1e-9)
    assert len(line) == 91 and line.endswith("\n")
    assert line[:15] == "     12.5000000" and line[30:45] == "     -0.2500000"
    assert line[60:75] == "*" * 15 and line[75:90] == "     -0.0000000"


def test_nxp_matches_grid_def():
    """NXP = NXV/NP - 1 with NXV = ceiling((NX+2)/NP) NP (rundir_unst/grid_def:29-31)."""
    assert PL.nxp(256, 16) == 16
    assert PL.nxp(1024, 16) == 64
    assert PL.nxp(64, 1) == 65
