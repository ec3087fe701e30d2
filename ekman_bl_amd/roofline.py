"""Algorithmic HBM bytes of the duct step (the byte contract, version 0).

One "pass" is one FP64 read or write of a full 3D field: 8 bytes per grid
point (a complex Fourier field of NX/2+1 modes occupies the same bytes).
F = 3 + N_TH transported fields.  A maximally fused plain FP64 substep must
move, across its global dependencies (explicit stage, y sweep, z sweep,
forward transform + divergence, projection):

    (5F+1) + 3F + 2F + (2F+1) + 9  passes per substep,

which over the three substeps of a step (no CF read at substep 1, no CF
write at substep 3) is 34F + 33 passes.  Each real (z,y) Poisson solve moves
15 plane passes for set-up, full multigrid and copies, plus 30.5 plane passes
per W(2,1) cycle; a substep solves the 2*(NX/3+1) real problems of the live
kx modes.  Counts are per grid point of N_pts = NX*(NY+2)*(NZ+2).
"""
from __future__ import annotations

PLANE_SETUP = 15.0
PLANE_PER_CYCLE = 30.5


def fields(p) -> int:
    return 3 + (1 if p.n_th else 0)


def mg_launch_bytes(p, nsolve: float, cycles_per_solve: float) -> float:
    """Algorithmic bytes of one batched multigrid launch."""
    plane = (p.ny + 2) * (p.nz + 2)
    return 8.0 * plane * nsolve * (PLANE_SETUP + PLANE_PER_CYCLE * cycles_per_solve)


def step_bytes(p, cycles_per_solve: float) -> float:
    """Algorithmic bytes of one full time step (three substeps)."""
    F = fields(p)
    npts = p.nx * (p.ny + 2) * (p.nz + 2)
    m_nz = p.nx // 3 + 1
    r = 2.0 * m_nz * (p.ny + 2) * (p.nz + 2) / npts
    return 8.0 * npts * ((34 * F + 33) + 3 * r * (PLANE_SETUP + PLANE_PER_CYCLE * cycles_per_solve))
