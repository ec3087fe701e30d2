"""LES_WALL_MODEL, WALL_MODEL_TYPE 3 (src/les_chan.F90:1914-2144), in the
oracle, against closed forms on a uniform rectilinear grid, where the face
metric CJOB_22 is the cell width in z and INT_JACOB the cell area, so the
row value of a Neumann row U(1) - U(0) = v is dy times the wall-normal
gradient the model prescribes:

- uniform U1 = u0, U3 = w0 at JSTART: the x means divide the sum over the NX
  points by NX-1 (MIN(NXP,NXP_L) for one rank), so U_1 = u0 NX/(NX-1),
  W_1 = w0 NX/(NX-1);
- U_STAR is the fixed point of the log law M = U_STAR (ln(Z U_STAR/NU)/0.41
  + 5.2) with M = |(U_1, W_1)| and Z = DYY(2,2) (thirty iterations);
- TAO_WALL = U_STAR^2 cos(alpha), TAO_WALL_1 = U_STAR^2 sin(alpha) with
  alpha = 20.6 degrees;
- the U3 row value is dy (w0/W_1) TAO_WALL/NU, the U1 one dy max(u0/W_1
  TAO_WALL, u0/U_1 TAO_WALL_1)/NU;
- on TIME_STEP 1 the U3 value is U3 at J = 2 and the U1 value is kept.
"""
import numpy as np
import pytest

from ekman_bl_amd.params import YMIN, ZMAX, ZMIN, base_params


def _wall_params(nx=16, ny=15, nz=15):
    p = base_params(nx, ny, nz, n_th=0)
    for c in range(3):
        p.u_bc[c][YMIN] = 3
    p.u_bc[0][ZMIN] = p.u_bc[0][ZMAX] = 1
    p.u_bc[1][ZMIN] = p.u_bc[1][ZMAX] = 1
    p.wall_model_type = 3
    p.nu = 1.0e-3
    return p


def test_wall_model_closed_forms(oracle_lib):
    O = oracle_lib
    p = _wall_params()
    dy, dz = 0.05, 0.25
    J = np.arange(p.ny + 2)[:, None] * np.ones((1, p.nz + 2))
    K = np.ones((p.ny + 2, 1)) * np.arange(p.nz + 2)[None, :]
    yp, xp = dy * J, dz * K
    o = O.Oracle(p, xp, yp)
    u0, w0 = 0.6, -0.9
    u1 = np.full((p.ny + 2, p.nz + 2, p.nx + 2), u0)
    u3 = np.full((p.ny + 2, p.nz + 2, p.nx + 2), w0)
    u3[2] = 0.123                                   # row J = 2 for the first-step branch
    out, info = o.wall_model(u1, u3, first=False)
    U_1, W_1, U_STAR, TAO, TAO1 = info
    f = p.nx / (p.nx - 1)
    assert U_1 == pytest.approx(u0 * f, rel=1e-14) and W_1 == pytest.approx(w0 * f, rel=1e-14)
    M = np.hypot(U_1, W_1)
    z = dy
    law = U_STAR * (np.log(z * U_STAR / p.nu) / 0.41 + float(np.float32(5.2)))
    assert law == pytest.approx(M, rel=1e-9)
    a = 20.6 * 3.1415 / 180.0          # the reference's angle, pi taken as 3.1415
    assert TAO == pytest.approx(U_STAR ** 2 * np.cos(a), rel=1e-6)
    assert TAO1 == pytest.approx(U_STAR ** 2 * np.sin(a), rel=1e-6)
    zs, ze, nx = 1, p.nz, p.nx
    w_rows = out[1, zs:ze + 1, :nx]
    u_rows = out[0, zs:ze + 1, :nx]
    np.testing.assert_allclose(w_rows, dy * (w0 / W_1) * TAO / p.nu, rtol=1e-7)
    np.testing.assert_allclose(u_rows, dy * max(u0 / W_1 * TAO, u0 / U_1 * TAO1) / p.nu, rtol=1e-7)
    assert not np.any(out[:, :, nx:]) and not np.any(out[:, 0]) and not np.any(out[:, ze + 1])
    # TIME_STEP 1: U3's value is U3 at J = 2, U1's is kept from before
    out1, _ = o.wall_model(u1 * 2.0, u3, first=True)
    np.testing.assert_array_equal(out1[1, zs:ze + 1, :nx], 0.123)
    np.testing.assert_array_equal(out1[0], out[0])


def test_wall_model_rejects_other_types(oracle_lib):
    """Types 1, 2 and 4 are not built (type 1 reads a U_STAR it never sets,
    type 2 a CU1X_WALL_MODEL it allocates and never assigns, type 4 the scalar's new
    value); the oracle refuses them."""
    O = oracle_lib
    p = _wall_params()
    from ekman_bl_amd.grid import tanh_grid
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, 2.0)
    for t in (1, 2, 4):
        p.wall_model_type = t
        with pytest.raises(Exception):
            O.Oracle(p, xp, yp)
