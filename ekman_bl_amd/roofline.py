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


# ---------------------------------------------------------------------------
# Consistent projection (proj_mode 1: FP64 cycles, 2: FP32 cycles inside FP64
# outer iterations).  Bytes per finest-plane point, element width w of the
# cycle (4 or 8); the coefficient tables are shared by all solves and not
# counted, the line factors are shared by the two solves of a mode (half
# counted per solve).  The multigrid part of a V(1,1) alternating-line cycle:
#   2 smoothing steps x 2 line directions: read r (w) + read/write e (2w)
#   + 3 factors / 2 solves (1.5w)                                   = 18w
#   residual (3w), restriction (1.5w), prolongation (2.25w)
#   -> 24.75w on the finest level, x 4/3 over all levels            = 33w
# Plain iterative refinement, per cycle:
#   FP64 residual and seed: read phi, b (16) + write r, clear e (2w)
#   phi += e: read e (w) + read/write phi (16)
#   => 32 + 36w.  Per solve once: read RHS_TMP, write b, clear phi, write
#   phi (32) and the final residual check (16 + 2w).
# GCR(1) outer iteration (proj_krylov 1), per cycle after the first:
#   q = A z: read z (w), r, Qp (16), write q (8)                    = 24 + w
#   update: read z (w), Zp, phi (16), write phi, Zp (16), clear z (w);
#   read q, Qp, r (24), write Qp, r (16), r in the cycle width (w)  = 72 + 4w
#   => 96 + 37w with the multigrid part.  Per solve once: read RHS_TMP,
#   write b, clear phi, write phi (32), the first residual: read phi, b,
#   write r (24) and its FP32 copy and the cleared e (2w) => 56 + 2w.
# ---------------------------------------------------------------------------
def tight_cycle_bytes_per_point(width: int, krylov: int = 0) -> float:
    return (96.0 + 37.0 * width) if krylov else (32.0 + 36.0 * width)


def tight_setup_bytes_per_point(width: int, krylov: int = 0) -> float:
    return (56.0 + 2.0 * width) if krylov else (48.0 + 2.0 * width)


def tight_launch_bytes(p, nsolve: int, cycles_total: float, width: int, krylov: int = 0) -> float:
    """Algorithmic bytes of one consistent-projection launch (all solves)."""
    plane = (p.ny + 2) * (p.nz + 2)
    return plane * (nsolve * tight_setup_bytes_per_point(width, krylov) +
                    cycles_total * tight_cycle_bytes_per_point(width, krylov))


def tight_step_bytes(p, cycles_per_projection: float, width: int, krylov: int = 0) -> float:
    """Algorithmic bytes of one time step with the consistent projection."""
    F = fields(p)
    npts = p.nx * (p.ny + 2) * (p.nz + 2)
    nsolve = 2 * (p.nx // 3 + 1)
    return 8.0 * npts * (34 * F + 33) + 3 * tight_launch_bytes(p, nsolve, cycles_per_projection, width, krylov)
