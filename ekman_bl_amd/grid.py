"""(z, y) grid of the duct: generation and GRID_YZ.dat input/output.

The reference reads the curvilinear grid from GRID_YZ.dat, one ``z y`` pair
per line, K fastest, for K = 0..NZ+1 and J = 0..NY+1
(src/diablo.F90:1120-1128).  Its shipped grid is a uniform z by a
tanh-stretched y, y_j = Ly * (1 - tanh(g*(1 - j/(N_y-1))) / tanh(g)); this
module generates that family.
"""
from __future__ import annotations

import numpy as np


def tanh_grid(ny: int, nz: int, ly: float, lz: float, gamma: float):
    """xpoint (z) and ypoint (y), each (NY+2, NZ+2) with K fastest."""
    n_y, n_z = ny + 2, nz + 2
    j = np.arange(n_y, dtype=np.float64)
    if gamma == 0.0:
        y = ly * j / (n_y - 1)
    else:
        y = ly * (1.0 - np.tanh(gamma * (1.0 - j / (n_y - 1))) / np.tanh(gamma))
    z = lz * np.arange(n_z, dtype=np.float64) / (n_z - 1)
    xpoint = np.broadcast_to(z[None, :], (n_y, n_z)).copy()
    ypoint = np.broadcast_to(y[:, None], (n_y, n_z)).copy()
    return xpoint, ypoint


def read_grid_yz(path: str, ny: int, nz: int):
    """Read a reference GRID_YZ.dat into (xpoint, ypoint), each (NY+2, NZ+2)."""
    a = np.loadtxt(path, dtype=np.float64)
    if a.shape != ((ny + 2) * (nz + 2), 2):
        raise ValueError(f"GRID_YZ.dat has {a.shape[0]} points, expected {(ny + 2) * (nz + 2)}")
    return a[:, 0].reshape(ny + 2, nz + 2).copy(), a[:, 1].reshape(ny + 2, nz + 2).copy()


def write_grid_yz(path: str, xpoint: np.ndarray, ypoint: np.ndarray) -> None:
    with open(path, "w") as f:
        for zv, yv in zip(xpoint.ravel(), ypoint.ravel()):
            f.write(f"{zv:.16e} {yv:.16e}\n")
