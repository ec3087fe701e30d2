"""Writes the small fixtures under tests/golden/ from the data files that ship
with the reference run directory rundir_unst/ (GRID_YZ.dat, input.dat,
grid_def).  Run in a container where the reference tree is mounted:

    python tests/golden/make_fixtures.py /root/reference

Outputs (committed):
  grid_yz_unst.json   rundir_unst/GRID_YZ.dat, losslessly: z(k) (the same on
                      every row), y(j) of column K=0, and the few (J,K,y)
                      entries whose printed y differs from its column value
                      in the last digits; plus the grid sizes from grid_def
  input_unst.dat      rundir_unst/input.dat, a data file, byte for byte
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def grid_def_sizes(path):
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from ekman_bl_amd.params import read_grid_def
    return read_grid_def(path)


def main(ref):
    run = os.path.join(ref, "rundir_unst")
    g = grid_def_sizes(os.path.join(run, "grid_def"))
    ny, nz = g["ny"], g["nz"]
    a = np.loadtxt(os.path.join(run, "GRID_YZ.dat"))
    assert a.shape == ((ny + 2) * (nz + 2), 2), a.shape
    zp = a[:, 0].reshape(ny + 2, nz + 2)
    yp = a[:, 1].reshape(ny + 2, nz + 2)
    z, y = zp[0, :].copy(), yp[:, 0].copy()
    assert np.array_equal(zp, np.broadcast_to(z[None, :], zp.shape))
    jj, kk = np.nonzero(yp != y[:, None])
    dev = [[int(j), int(k), float(yp[j, k])] for j, k in zip(jj, kk)]
    rebuilt = np.broadcast_to(y[:, None], yp.shape).copy()
    for j, k, v in dev:
        rebuilt[j, k] = v
    assert np.array_equal(rebuilt, yp)
    keys = ("nx", "ny", "nz", "n_th", "iixp", "jjyq", "iiex", "jjey", "np", "ny_s_t", "nz_s_r", "nz_s_l")
    fix = {"source": "rundir_unst/GRID_YZ.dat and rundir_unst/grid_def",
           "grid_def": {k: g[k] for k in keys if k in g},
           "y": [float(v) for v in y], "z": [float(v) for v in z], "y_exceptions": dev}
    with open(os.path.join(HERE, "grid_yz_unst.json"), "w") as f:
        json.dump(fix, f, indent=0)
    shutil.copyfile(os.path.join(run, "input.dat"), os.path.join(HERE, "input_unst.dat"))
    print("wrote grid_yz_unst.json, input_unst.dat")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
