"""The reference's plane, 3D and bulk-velocity output of SAVE_STATS_CURVI
(src/flow_statistics.F90:76-97, 379-405, 571-810, 2004-2629), written from
data the solver gathers on the device: the physical planes of ekm_get_plane,
the physical fields of ekm_get_physical and the x means of ekm_get_xstats.

Files are Fortran unformatted sequential, record for record as the reference
writes them (4-byte little-endian markers, REAL*8 data, default INTEGER
headers, arrays in Fortran order: the first index fastest).

The reference assembles the "combined" x-y and x-z planes from the x slabs
of its NP ranks on rank 0, and NXP (the last x index of a rank,
``NXP = NXV/NP - 1``, ``NXV = ceiling((NX+2)/NP) NP``, rundir_unst/grid_def:
28-31) enters the headers, the z-y plane positions and the time_bulk probe.
One device holds every x point here; `nranks` names the NP whose files are
reproduced (16, the shipped grid_def's, by default).  The global plane is the
same for every NP.

Two points where the reference leaves values undefined are written as zeros:
the x-y vorticity omg_z outside I = 1..NX-1, J = 1..NY (a local array the
reference never initialises there), and NX+1 x points of a one-rank run
(with NP = 1 the reference copies NX+2 points, the two padding entries
included, into its (0:NX) plane; the combined writers warn that they do not
support that case).
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

from .restart import write_record

XZ_ROWS = (2, 13, 29, 44, 51, 70, 98, 109)   # src/flow_statistics.F90:758-802


def nxp(nx: int, nranks: int) -> int:
    """NXP of grid_def: the last local x index of a rank."""
    nxv = math.ceil((nx + 2) / nranks) * nranks
    return nxv // nranks - 1


def _seq_sum(a: np.ndarray, axis: int) -> np.ndarray:
    """Left-to-right sum (Fortran SUM over a section), not numpy's pairwise sum."""
    return np.cumsum(a, axis=axis).take(-1, axis=axis)


def _cdiff(h1, h2, up, um):
    """The writers' one-sided-weight central difference b u(+1) + c u(-1) with
    b = h2/h1/(h1+h2), c = -h1/h2/(h1+h2); the centre term of the
    non-uniform three-point formula is absent, as in the reference
    (src/flow_statistics.F90:2283-2290, 2583-2590)."""
    b = h2 / h1 / (h1 + h2)
    c = -h1 / h2 / (h1 + h2)
    return b * up + c * um


def xy_combined(time: float, k: int, ypoint: np.ndarray, dx: float, planes: np.ndarray, nxp_val: int):
    """plane_XY_binary_combined (src/flow_statistics.F90:2091-2303) on the
    physical x-y planes at K = k (planes: U1, U2, U3, P, TH, each (NY+2, NX)).
    Returns the three records as lists of arrays / header tuple."""
    nf, ny2, nx = planes.shape
    ny = ny2 - 2
    tot = np.zeros((5, ny2, nx + 1))
    tot[:, :, :nx] = planes
    tot[:, :, nx] = planes[:, :, 0]                     # corner update (:2201-2206)
    xt = np.broadcast_to((np.arange(nx + 1) * dx)[None, :], (ny2, nx + 1)).copy()
    yt = np.broadcast_to(ypoint[:, k][:, None], (ny2, nx + 1)).copy()
    omg = np.zeros((ny2, nx + 1))
    J, I = slice(1, ny + 1), slice(1, nx)
    dudy = _cdiff(yt[2:ny + 2, I] - yt[J, I], yt[J, I] - yt[0:ny, I], tot[0, 2:ny + 2, I], tot[0, 0:ny, I])
    dvdx = _cdiff(xt[J, 2:nx + 1] - xt[J, I], xt[J, I] - xt[J, 0:nx - 1], tot[1, J, 2:nx + 1], tot[1, J, 0:nx - 1])
    omg[J, I] = dvdx - dudy
    head = (time, nx, ny, k, nxp_val)
    return head, [yt, xt], [tot[0], tot[1], tot[2], tot[3], tot[4], omg]


def xz_combined(time: float, j: int, xpoint: np.ndarray, dx: float, planes: np.ndarray, nxp_val: int):
    """plane_XZ_binary_combined (src/flow_statistics.F90:2374-2629) on the
    physical x-z planes at J = j (planes: U1, U2, U3, P, TH, each (NZ+2, NX)),
    over K = 1..NZ.  The U1 fluctuation subtracts the z mean of x point 0 at
    every x point: the reference indexes that mean with its stale rank-loop
    variable I (0 after the receive loop) where the other fields use II
    (:2603)."""
    _, nz2, nx = planes.shape
    nz = nz2 - 2
    tot = planes[:, 1:nz + 1, :].copy()                  # (5, NZ, NX): K = 1..NZ
    xt = np.broadcast_to((np.arange(nx) * dx)[None, :], (nz, nx)).copy()
    zt = np.broadcast_to(xpoint[j, 1:nz + 1][:, None], (nz, nx)).copy()

    def vort(u1, u3):
        o = np.zeros((nz, nx))
        K, I = slice(1, nz - 1), slice(1, nx - 1)       # JJ = 2..NZ-1, II = 1..NX-2
        dudz = _cdiff(zt[2:nz, I] - zt[K, I], zt[K, I] - zt[0:nz - 2, I], u1[2:nz, I], u1[0:nz - 2, I])
        dwdx = _cdiff(xt[K, 2:nx] - xt[K, I], xt[K, I] - xt[K, 0:nx - 2], u3[K, 2:nx], u3[K, 0:nx - 2])
        o[K, I] = dudz - dwdx
        o[:, 0] = o[:, nx - 2]
        o[:, nx - 1] = o[:, 1]
        o[nz - 1, :] = o[1, :]
        o[0, :] = o[nz - 2, :]
        return o

    omg = vort(tot[0], tot[2])
    means = _seq_sum(tot, axis=1) / float(nz)           # (5, NX): z mean of every x point
    fl = tot - means[:, None, :]
    fl[0] = tot[0] - means[0, 0]
    omg_p = vort(fl[0], fl[2])
    head = (time, nx, nz, j, nxp_val)
    return head, [zt, xt], list(tot), list(fl) + [omg, omg_p]


def _f64(*arrays) -> bytes:
    return b"".join(np.ascontiguousarray(a, dtype="<f8").tobytes() for a in arrays)


def write_xy_combined(path: str, *args) -> None:
    head, grid, data = xy_combined(*args)
    with open(path, "wb") as f:
        write_record(f, struct.pack("<d4i", *head))
        write_record(f, _f64(*grid))
        write_record(f, _f64(*data))


def write_xz_combined(path: str, *args) -> None:
    head, grid, data, fluct = xz_combined(*args)
    with open(path, "wb") as f:
        write_record(f, struct.pack("<d4i", *head))
        write_record(f, _f64(*grid))
        write_record(f, _f64(*data))
        write_record(f, _f64(*fluct))


def write_zy(path: str, time: float, xpoint: np.ndarray, ypoint: np.ndarray, planes: np.ndarray) -> None:
    """plane_ZY_binary (src/flow_statistics.F90:2310-2336): one record of
    TIME, xpoint, ypoint and U3, U2, U1, P, TH on the z-y plane (each
    (NY+2, NZ+2))."""
    with open(path, "wb") as f:
        write_record(f, struct.pack("<d", time) + _f64(xpoint, ypoint, planes[2], planes[1], planes[0],
                                                       planes[3], planes[4]))


def write_3d(path: str, nx: int, xpoint: np.ndarray, ypoint: np.ndarray, dx: float, fields, th_bar=None) -> None:
    """plot_3D (src/flow_statistics.F90:2004-2040) for one rank holding every
    x point (MIN(NXP, NXP_L) = NX-1): records (xpoint, ypoint, dx, U3), U2,
    U1, TH + TH_BAR, P, each field over x points 0..NX-1.  fields: physical
    U1, U2, U3, P, TH, each (NY+2, NZ+2, >= NX); TH None for N_TH = 0 (an
    empty record); th_bar (NY+2, NZ+2) or None for a zero background."""
    u1, u2, u3, p, th = (None if a is None else a[:, :, :nx] for a in fields)
    with open(path, "wb") as f:
        write_record(f, _f64(xpoint, ypoint) + struct.pack("<d", dx) + _f64(u3))
        write_record(f, _f64(u2))
        write_record(f, _f64(u1))
        if th is None:
            write_record(f, b"")
        else:
            tb = np.zeros(th.shape[:2]) if th_bar is None else th_bar
            write_record(f, _f64(th + tb[:, :, None]))
        write_record(f, _f64(p))


def bulk(u3_mean: np.ndarray, u1_mean: np.ndarray):
    """UBULK and VBULK (src/flow_statistics.F90:83-95): the corner averages of
    the mode-0 U3 and U1 summed over K = 1..NZ, J = 1..NY (J outer), over
    REAL(NY*NZ) in single precision.  Inputs (NY+2, NZ+2)."""
    ny, nz = u3_mean.shape[0] - 2, u3_mean.shape[1] - 2

    def one(m):
        c = 0.25 * (m[1:ny + 1, 1:nz + 1] + m[1:ny + 1, 0:nz] + m[0:ny, 0:nz] + m[0:ny, 1:nz + 1])
        return float(np.cumsum(c.ravel())[-1]) / float(np.float32(ny * nz))

    return one(u3_mean), one(u1_mean)


def _f15_7(v: float) -> str:
    s = "%15.7f" % v
    return "*" * 15 if len(s) > 15 else s


def time_bulk_line(time: float, delta_t: float, ubulk: float, vbulk: float, u3_probe: float, u1_probe: float) -> str:
    """One line of plane_data/time_bulk.txt, FORMAT(6f15.7) (:402-404)."""
    return "".join(_f15_7(v) for v in (time, delta_t, ubulk, vbulk, u3_probe, u1_probe)) + "\n"


def _num4(n: int) -> str:
    return "%04d" % (n % 10000)


def save_stats_output(solver, run_dir: str, time_step: int, save_stats_int: int, xpoint: np.ndarray,
                      ypoint: np.ndarray, nranks: int = 16, save_3d: bool = False, span_xy: bool = True,
                      span_zy: bool = True, span_xz: bool = True, th_bar=None) -> list:
    """The output block of SAVE_STATS_CURVI for the solver's current state:
    plane_data/time_bulk.txt (appended, :379-405), plane_3D/data_3d_NNNN.pln_001
    when save_3d (:571-596), xy_plane/span_NNNN_K_kkkk.pln at
    K = (NZ+1)/4, (NZ+1)/2, 3(NZ+1)/4 (:598-647), zy_plane/spanM_NNNN_.pln at
    the reference's three x positions (:649-705) and
    xz_plane/span_NNNN_J_jjjj.pln at J = 2, 13, 29, 44, 51, 70, 98, 109
    (:707-802; rows beyond NY+1 are skipped), NNNN = TIME_STEP/SAVE_STATS_INT.
    The reference's defaults are SAVE_3D false and the three plane kinds on
    (:38-42).  Returns the paths written."""
    p = solver.p
    nx, ny, nz = p.nx, p.ny, p.nz
    dx = p.lx / nx                                          # DX(I) = LX/NX (src/diablo.F90:409)
    d = solver.diag()
    xp_ = nxp(nx, nranks)
    num = _num4(time_step // save_stats_int)
    out = []

    def mk(sub):
        os.makedirs(os.path.join(run_dir, sub), exist_ok=True)

    st = solver.x_stats()
    ub, vb = bulk(st["u3"], st["u1"])
    probe = solver.plane(0, 2)                              # U3X(NXP, 2, NY), U1X(NXP, 2, NY) of rank 0
    ip = min(xp_, nx - 1)
    mk("plane_data")
    path = os.path.join(run_dir, "plane_data", "time_bulk.txt")
    with open(path, "a") as f:
        f.write(time_bulk_line(d.time, d.delta_t, ub, vb, probe[2, ny, ip], probe[0, ny, ip]))
    out.append(path)
    if save_3d:
        mk("plane_3D")
        path = os.path.join(run_dir, "plane_3D", f"data_3d_{num}.pln_001")
        fields = [solver.physical(f) for f in range(4)] + [solver.physical(4) if p.n_th else None]
        write_3d(path, nx, xpoint, ypoint, dx, fields, th_bar)
        out.append(path)
    if span_xy:
        mk("xy_plane")
        for k in ((nz + 1) // 4, (nz + 1) // 2, (3 * (nz + 1)) // 4):
            path = os.path.join(run_dir, "xy_plane", f"span_{num}_K_{k:04d}.pln")
            write_xy_combined(path, d.time, k, ypoint, dx, solver.plane(0, k), xp_)
            out.append(path)
    if span_zy:
        mk("zy_plane")
        w = xp_ + 1
        # (rank, local index) pairs of :676-698; 4/3 is integer division there
        for name, rank, loc in (("span1", nx // w // 4, (nx // 4) % w), ("span2", nx // w // 2, (nx // 2) % w),
                                ("span3", nx // w // (4 // 3), (4 * nx // 3) % w)):
            i = rank * w + loc
            if rank >= nranks or i >= nx:
                continue
            path = os.path.join(run_dir, "zy_plane", f"{name}_{num}_.pln")
            write_zy(path, d.time, xpoint, ypoint, solver.plane(1, i))
            out.append(path)
    if span_xz:
        mk("xz_plane")
        for j in XZ_ROWS:
            if j > ny + 1:
                continue
            path = os.path.join(run_dir, "xz_plane", f"span_{num}_J_{j:04d}.pln")
            write_xz_combined(path, d.time, j, xpoint, dx, solver.plane(2, j), xp_)
            out.append(path)
    return out
