"""Restart files in the reference's format: SAVE_FLOW, BACKUP_FLOW and
READ_FLOW (src/diablo.F90:753-1072).

Each file is Fortran unformatted sequential (4-byte little-endian record
markers before and after every record) with two records:

1. ``NX, NY, NZ, TIME, TIME_STEP`` -- three default integers, a REAL*8 and a
   default integer (24 bytes);
2. the Fourier fields as COMPLEX*16 in Fortran order, kx fastest, then z,
   then y: ``CU1X(0:NX2P, 0:NZ+1, 0:NY+1)``, then CU2X, CU3X, CPX in the
   velocity file (src/diablo.F90:815-819), ``CTHX(..., N)`` in the file of
   scalar N (:822-825).

File names (rank r, 1-based ``I = r + 1``; scalar N): ``diablo_III.saved``
and ``diablo_thNN_III.saved`` under ``last_saved/`` (SAVE_FLOW with FINAL
false, the files READ_FLOW reads), ``.res`` for SAVE_FLOW(FINAL), and
``backup/diablo_III.saved_MMMM`` with ``MMMM = TIME_STEP / BACKUP_FLOW_INT``
for BACKUP_FLOW.

With NP ranks every rank writes the kx modes it owns: NX2P + 1 of them per
rank, ``NX2V = ceiling((NX+2)/NP / 2) * NP``, ``NX2P = NX2V/NP - 1``
(rundir_unst/grid_def:31-37), rank r holding modes ``r (NX2P+1)`` upwards.
read_flow merges the ranks' files and keeps modes 0..NX/2.

Records longer than 2**31 - 1 bytes (config 5's field record is 9.7 GB) are
split into subrecords whose leading and trailing markers carry the length
with the sign bit set on every subrecord but the last (leading marker) and
every subrecord but the first (trailing marker), the scheme gfortran and the
Intel compiler share.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass

import numpy as np

MAX_SUBRECORD = 2**31 - 1


# ---------------------------------------------------------------------------
# Fortran unformatted sequential records
# ---------------------------------------------------------------------------
def write_record(f, payload: bytes | memoryview, max_sub: int = MAX_SUBRECORD) -> None:
    """One logical record, split into subrecords of at most max_sub bytes."""
    mv = memoryview(payload).cast("B")
    n = len(mv)
    if n == 0:
        f.write(struct.pack("<ii", 0, 0))
        return
    pos = 0
    first = True
    while pos < n:
        k = min(max_sub, n - pos)
        last = pos + k == n
        lead = k if last else -k
        trail = k if first else -k
        f.write(struct.pack("<i", lead))
        f.write(mv[pos:pos + k])
        f.write(struct.pack("<i", trail))
        pos += k
        first = False


def read_record(f) -> bytes:
    """The payload of the next logical record (subrecords joined)."""
    parts = []
    while True:
        hdr = f.read(4)
        if len(hdr) < 4:
            raise EOFError("end of file inside a record")
        lead = struct.unpack("<i", hdr)[0]
        k = abs(lead)
        data = f.read(k)
        if len(data) != k:
            raise EOFError("truncated record")
        tail = f.read(4)
        if len(tail) < 4 or abs(struct.unpack("<i", tail)[0]) != k:
            raise ValueError("record markers do not match")
        parts.append(data)
        if lead >= 0:
            return b"".join(parts)


def _read_record_into(f, out: np.ndarray) -> None:
    """Read the next record straight into out's bytes (no copy of the whole
    record in memory, which matters for multi-GB field records)."""
    buf = memoryview(out.reshape(-1).view(np.uint8))
    pos = 0
    while True:
        hdr = f.read(4)
        if len(hdr) < 4:
            raise EOFError("end of file inside a record")
        lead = struct.unpack("<i", hdr)[0]
        k = abs(lead)
        if pos + k > len(buf):
            raise ValueError(f"record is longer than the expected {len(buf)} bytes")
        if f.readinto(buf[pos:pos + k]) != k:
            raise EOFError("truncated record")
        pos += k
        tail = f.read(4)
        if len(tail) < 4 or abs(struct.unpack("<i", tail)[0]) != k:
            raise ValueError("record markers do not match")
        if lead >= 0:
            break
    if pos != len(buf):
        raise ValueError(f"record holds {pos} bytes, expected {len(buf)}")


# ---------------------------------------------------------------------------
# the reference's restart files
# ---------------------------------------------------------------------------
HEADER = struct.Struct("<iiidi")  # NX, NY, NZ, TIME, TIME_STEP


@dataclass
class Flow:
    """A restart state: Fourier fields shaped (NY+2, NZ+2, NX/2+1) complex128
    (C order, so kx is fastest as in the file), TIME and TIME_STEP."""
    nx: int
    ny: int
    nz: int
    time: float
    time_step: int
    cu1: np.ndarray
    cu2: np.ndarray
    cu3: np.ndarray
    cp: np.ndarray
    cth: list  # one array per scalar


def modes_per_rank(nx: int, nranks: int) -> int:
    """NX2P + 1 of rundir_unst/grid_def:31-37 for NP = nranks."""
    nxv = -(-(nx + 2) // nranks) * nranks
    nx2v = -(-(nxv // 2) // nranks) * nranks
    return nx2v // nranks


def flow_name(rank: int = 0, kind: str = "saved", scalar: int | None = None, backup_index: int | None = None) -> str:
    """File name (relative to the run directory) of rank `rank` (0-based).
    kind: "saved" (SAVE_FLOW(.FALSE.), READ_FLOW), "res" (SAVE_FLOW(.TRUE.))
    or "backup" (BACKUP_FLOW, with backup_index = TIME_STEP/BACKUP_FLOW_INT)."""
    i = rank + 1
    stem = f"diablo_{i % 1000:03d}" if scalar is None else f"diablo_th{scalar % 100:02d}_{i % 1000:03d}"
    if kind == "backup":
        if backup_index is None:
            raise ValueError("backup files need backup_index")
        return os.path.join("backup", f"{stem}.saved_{backup_index % 10000:04d}")
    if kind not in ("saved", "res"):
        raise ValueError(f"unknown restart kind {kind!r}")
    return os.path.join("last_saved", f"{stem}.{kind}")


def _write_file(path: str, nx: int, ny: int, nz: int, time: float, time_step: int, fields) -> None:
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "wb") as f:
        write_record(f, HEADER.pack(nx, ny, nz, float(time), int(time_step)))
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in fields]
        if len(arrs) == 1:
            write_record(f, memoryview(arrs[0]).cast("B"))
        else:
            write_record(f, b"".join(memoryview(a).cast("B") for a in arrs))


def save_flow(run_dir: str, flow: Flow, kind: str = "saved", backup_index: int | None = None,
              nranks: int = 1) -> list[str]:
    """Write the velocity/pressure file and one file per scalar as SAVE_FLOW or
    BACKUP_FLOW would with NP = nranks; returns the paths written."""
    nxc = flow.nx // 2 + 1
    for a in (flow.cu1, flow.cu2, flow.cu3, flow.cp, *flow.cth):
        if a.shape != (flow.ny + 2, flow.nz + 2, nxc):
            raise ValueError(f"field shape {a.shape} != {(flow.ny + 2, flow.nz + 2, nxc)}")
    per = modes_per_rank(flow.nx, nranks)
    paths = []
    for r in range(nranks):
        lo = r * per

        def part(a):
            out = np.zeros((flow.ny + 2, flow.nz + 2, per), dtype=np.complex128)
            hi = min(lo + per, nxc)
            if hi > lo:
                out[..., :hi - lo] = a[..., lo:hi]
            return out

        path = os.path.join(run_dir, flow_name(r, kind, None, backup_index))
        _write_file(path, flow.nx, flow.ny, flow.nz, flow.time, flow.time_step,
                    [part(flow.cu1), part(flow.cu2), part(flow.cu3), part(flow.cp)])
        paths.append(path)
        for n, th in enumerate(flow.cth, start=1):
            tpath = os.path.join(run_dir, flow_name(r, kind, n, backup_index))
            _write_file(tpath, flow.nx, flow.ny, flow.nz, flow.time, flow.time_step, [part(th)])
            paths.append(tpath)
    return paths


def _read_header(f, nx: int | None, ny: int | None, nz: int | None):
    h = read_record(f)
    if len(h) != HEADER.size:
        raise ValueError(f"header record of {len(h)} bytes, expected {HEADER.size}")
    fnx, fny, fnz, t, ts = HEADER.unpack(h)
    if (nx, ny, nz) != (None, None, None) and (fnx, fny, fnz) != (nx, ny, nz):
        # READ_FLOW's check (src/diablo.F90:955-958)
        raise ValueError(f"old flowfield wrong dimensions: file {fnx, fny, fnz}, run {nx, ny, nz}")
    return fnx, fny, fnz, t, ts


def read_flow(run_dir: str, nx: int | None = None, ny: int | None = None, nz: int | None = None,
              n_th: int = 0, nranks: int = 1, kind: str = "saved", backup_index: int | None = None) -> Flow:
    """Read the files READ_FLOW reads (or the .res / backup files), merging
    nranks ranks' kx modes; the grid size must match nx, ny, nz when given.
    TIME and TIME_STEP come from the velocity file (READ_FLOW reads the scalar
    files' headers over them; SAVE_FLOW writes the same values to both)."""
    fields = None
    per = None
    for r in range(nranks):
        path = os.path.join(run_dir, flow_name(r, kind, None, backup_index))
        with open(path, "rb") as f:
            fnx, fny, fnz, t, ts = _read_header(f, nx, ny, nz)
            if fields is None:
                nx, ny, nz, time, time_step = fnx, fny, fnz, t, ts
                per = modes_per_rank(nx, nranks)
                nxc = nx // 2 + 1
                fields = [np.zeros((ny + 2, nz + 2, nxc), dtype=np.complex128) for _ in range(4 + n_th)]
            block = np.empty((4, ny + 2, nz + 2, per), dtype=np.complex128)
            _read_record_into(f, block)
        lo = r * per
        hi = min(lo + per, nx // 2 + 1)
        for q in range(4):
            if hi > lo:
                fields[q][..., lo:hi] = block[q][..., :hi - lo]
        for n in range(1, n_th + 1):
            tpath = os.path.join(run_dir, flow_name(r, kind, n, backup_index))
            with open(tpath, "rb") as f:
                _read_header(f, nx, ny, nz)
                tb = np.empty((ny + 2, nz + 2, per), dtype=np.complex128)
                _read_record_into(f, tb)
            if hi > lo:
                fields[3 + n][..., lo:hi] = tb[..., :hi - lo]
    return Flow(nx, ny, nz, time, time_step, fields[0], fields[1], fields[2], fields[3], fields[4:])
