"""The statistics output of SAVE_STATS_CURVI on the GPU (through the C ABI)
against the oracle, after one step of a small stratified case (the
256x193x257 case's physics on 32 x 33 x 33 points, sheared grid):

- ekm_get_omega_x, the mean streamwise vorticity (src/flow_statistics.F90:
  301-327), to 1e-12 of its scale;
- ekm_get_physical and the three plane kinds of ekm_get_plane against the
  oracle's physical fields, to 1e-12 of each field's scale;
- the files of ekman_bl_amd.planes.save_stats_output (time_bulk.txt, the
  x-y, z-y and x-z planes and the 3D file) against the same writers fed
  with the oracle's fields, record by record to 1e-12.
"""
import os

import numpy as np
import pytest

from ekman_bl_amd import planes as PL
from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import config, mg_factors
from ekman_bl_amd.restart import read_record

pytestmark = pytest.mark.gpu


def _pair(oracle_lib):
    cfg = config(2)
    p = cfg.params
    p.nx, p.ny, p.nz = 32, 31, 31
    p.mg_ixp, p.mg_iex = mg_factors(p.nz + 2)
    p.mg_jyq, p.mg_jey = mg_factors(p.ny + 2)
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    xp = xp + 0.05 * np.sin(0.3 * yp)
    ic = ekman_initial(p, yp, cfg.seed, 0.05, 0.01)
    o = oracle_lib.Oracle(p, xp, yp)
    s = S.DuctSolver(p, xp, yp)
    o.set_initial_physical(*ic)
    s.set_initial_physical(*ic)
    o.time_step(1)
    s.time_step(1)
    return p, xp, yp, o, s


def _close(a, b, what):
    scale = max(np.max(np.abs(b)), 1e-300)
    err = np.max(np.abs(a - b)) / scale
    print(f"[output] {what}: {err:.2e} of {scale:.3e}", flush=True)
    assert err < 1e-12, what


def test_omega_x_and_planes_match_oracle(oracle_lib):
    p, xp, yp, o, s = _pair(oracle_lib)
    _close(s.omega_x(), o.omega_x(), "omega_x")
    phys = [o.physical(f) for f in range(5)]
    for f in range(5):
        _close(s.physical(f)[:, :, :p.nx], phys[f][:, :, :p.nx], f"physical {f}")
    for k in (0, 7, p.nz + 1):
        pl = s.plane(0, k)
        for f in range(5):
            _close(pl[f], phys[f][:, k, :p.nx], f"x-y plane K={k} field {f}")
    for i in (0, 5, p.nx - 1):
        pl = s.plane(1, i)
        for f in range(5):
            _close(pl[f], phys[f][:, :, i], f"z-y plane I={i} field {f}")
    for j in (0, 13, p.ny + 1):
        pl = s.plane(2, j)
        for f in range(5):
            _close(pl[f], phys[f][j, :, :p.nx], f"x-z plane J={j} field {f}")
    with pytest.raises(S.EkmError):
        s.plane(1, p.nx)
    s.close()


class _OracleSource:
    """The accessors save_stats_output reads, served from the oracle."""

    def __init__(self, p, o):
        self.p, self.o = p, o
        self.phys = [o.physical(f) for f in range(5)]

    def diag(self):
        return self.o.diag()

    def x_stats(self):
        return self.o.x_stats()

    def plane(self, axis, index):
        nx = self.p.nx
        if axis == 0:
            return np.stack([f[:, index, :nx] for f in self.phys])
        if axis == 1:
            return np.stack([f[:, :, index] for f in self.phys])
        return np.stack([f[index, :, :nx] for f in self.phys])

    def physical(self, f):
        return self.phys[f]


def _all_records(path):
    out = []
    with open(path, "rb") as f:
        while True:
            try:
                out.append(read_record(f))
            except EOFError:
                return out


def test_save_stats_output_matches_oracle(oracle_lib, tmp_path):
    p, xp, yp, o, s = _pair(oracle_lib)
    gdir, odir = str(tmp_path / "gpu"), str(tmp_path / "oracle")
    got = PL.save_stats_output(s, gdir, 40, 10, xp, yp, save_3d=True)
    want = PL.save_stats_output(_OracleSource(p, o), odir, 40, 10, xp, yp, save_3d=True)
    names = sorted(os.path.relpath(x, gdir) for x in got)
    assert names == sorted(os.path.relpath(x, odir) for x in want)
    assert "xy_plane/span_0004_K_0008.pln" in names and "plane_3D/data_3d_0004.pln_001" in names
    assert "xz_plane/span_0004_J_0029.pln" in names and "xz_plane/span_0004_J_0044.pln" not in names
    for rel in names:
        a, b = os.path.join(gdir, rel), os.path.join(odir, rel)
        if rel.endswith(".txt"):
            va = np.array(open(a).read().split(), float)
            vb = np.array(open(b).read().split(), float)
            assert np.allclose(va, vb, rtol=1e-6, atol=2e-7), rel
            continue
        ra, rb = _all_records(a), _all_records(b)
        assert [len(x) for x in ra] == [len(x) for x in rb], rel
        for n, (x, y) in enumerate(zip(ra, rb)):
            if rel.startswith(("xy", "xz")) and n == 0:
                assert x[8:] == y[8:], rel            # NX, NY or NZ, K or J, NXP
                _close(np.frombuffer(x[:8], "<f8"), np.frombuffer(y[:8], "<f8"), f"{rel} TIME")
                continue
            _close(np.frombuffer(x, "<f8"), np.frombuffer(y, "<f8"), f"{rel} record {n}")
    s.close()
