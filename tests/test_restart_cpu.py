"""Restart files in the reference's SAVE_FLOW / READ_FLOW format
(src/diablo.F90:753-1072), ekman_bl_amd/restart.py.

The format is pinned against a Fortran runtime: tests/fortran/restart_io.f90
writes and reads the two records with the reference's WRITE / READ statement
form (implied-DO lists over I, K, J of COMPLEX*16, header NX, NY, NZ, TIME,
TIME_STEP), built with amdflang; Python must read what it writes and write
what it reads.  Rank splitting, scalar files, file names, subrecords and the
dimension check are covered on the Python side.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from ekman_bl_amd import restart as R

HERE = os.path.dirname(os.path.abspath(__file__))
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def _expected(nx, ny, nz, f):
    """The closed formula of restart_io.f90 for field f (1..4), (NY+2, NZ+2, NX/2+1)."""
    j, k, i = np.meshgrid(np.arange(ny + 2), np.arange(nz + 2), np.arange(nx // 2 + 1), indexing="ij")
    return (i + 1000.0 * k + 1.0e6 * j + 0.25 * f) - 1j * (f + 0.5 * i - 7.0 * k + 3.0 * j)


@pytest.fixture(scope="module")
def fortran_io(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("rio")
    exe = str(d / "restart_io")
    r = subprocess.run([FLANG, "-O1", "-o", exe, os.path.join(HERE, "fortran", "restart_io.f90")],
                       capture_output=True, text=True, cwd=str(d), timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("nx,ny,nz", [(8, 3, 5), (16, 7, 9)])
def test_fortran_written_file_is_read(nx, ny, nz, fortran_io, tmp_path):
    path = tmp_path / R.flow_name(0, "saved")
    path.parent.mkdir(parents=True)
    r = subprocess.run([fortran_io, "w", str(path), str(nx), str(ny), str(nz)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fl = R.read_flow(str(tmp_path), nx, ny, nz)
    assert (fl.time, fl.time_step) == (1.25, 42)
    for f, a in enumerate((fl.cu1, fl.cu2, fl.cu3, fl.cp), start=1):
        assert np.array_equal(a, _expected(nx, ny, nz, f)), f


@pytest.mark.parametrize("nx,ny,nz", [(8, 3, 5), (16, 7, 9)])
def test_python_written_file_reads_in_fortran(nx, ny, nz, fortran_io, tmp_path):
    fl = R.Flow(nx, ny, nz, 1.25, 42, *[_expected(nx, ny, nz, f) for f in range(1, 5)], [])
    (path,) = R.save_flow(str(tmp_path), fl)
    assert path.endswith(os.path.join("last_saved", "diablo_001.saved"))
    r = subprocess.run([fortran_io, "r", path, str(nx), str(ny), str(nz)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "ok" in r.stdout, r.stdout + r.stderr


def _random_flow(nx, ny, nz, n_th, seed=3):
    rng = np.random.default_rng(seed)
    shape = (ny + 2, nz + 2, nx // 2 + 1)
    f = [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(4 + n_th)]
    return R.Flow(nx, ny, nz, 12.5, 300, f[0], f[1], f[2], f[3], f[4:])


@pytest.mark.parametrize("nranks", [1, 2, 3, 5])
def test_round_trip_with_ranks_and_scalar(nranks, tmp_path):
    """Every rank's file carries NX2P+1 modes (rundir_unst/grid_def:31-37),
    the last rank zero-padded; reading merges them back."""
    fl = _random_flow(32, 5, 7, 1)
    paths = R.save_flow(str(tmp_path), fl, nranks=nranks)
    assert len(paths) == 2 * nranks
    per = R.modes_per_rank(32, nranks)
    assert per * nranks >= 17
    for r in range(nranks):
        sz = os.path.getsize(tmp_path / R.flow_name(r, "saved"))
        assert sz == (24 + 8) + (4 * 7 * 9 * per * 16 + 8)
        assert os.path.exists(tmp_path / R.flow_name(r, "saved", scalar=1))
    back = R.read_flow(str(tmp_path), 32, 5, 7, n_th=1, nranks=nranks)
    assert (back.time, back.time_step) == (12.5, 300)
    for a, b in zip((fl.cu1, fl.cu2, fl.cu3, fl.cp, fl.cth[0]), (back.cu1, back.cu2, back.cu3, back.cp, back.cth[0])):
        assert np.array_equal(a, b)


def test_modes_per_rank_matches_grid_def():
    """NX2P + 1 at NP = 1 is NX/2 + 1 (the whole Fourier field); the ranks
    together always cover modes 0..NX/2."""
    for nx in (64, 256, 512, 1024):
        assert R.modes_per_rank(nx, 1) == nx // 2 + 1
        for npr in (2, 4, 8, 16):
            assert R.modes_per_rank(nx, npr) * npr >= nx // 2 + 1


def test_file_names():
    assert R.flow_name(0, "saved") == os.path.join("last_saved", "diablo_001.saved")
    assert R.flow_name(11, "res") == os.path.join("last_saved", "diablo_012.res")
    assert R.flow_name(0, "saved", scalar=1) == os.path.join("last_saved", "diablo_th01_001.saved")
    assert R.flow_name(2, "backup", backup_index=7) == os.path.join("backup", "diablo_003.saved_0007")
    assert R.flow_name(0, "backup", scalar=2, backup_index=12) == os.path.join("backup", "diablo_th02_001.saved_0012")


def test_subrecords(tmp_path):
    """A record longer than the subrecord limit is split: leading markers
    negative on every subrecord but the last, trailing markers negative on
    every subrecord but the first; the reader joins them."""
    payload = bytes(range(23))
    p = tmp_path / "r.bin"
    with open(p, "wb") as f:
        R.write_record(f, payload, max_sub=10)
        R.write_record(f, b"xy", max_sub=10)
    raw = open(p, "rb").read()
    leads, pos = [], 0
    for k in (10, 10, 3):
        lead, = struct.unpack("<i", raw[pos:pos + 4])
        trail, = struct.unpack("<i", raw[pos + 4 + k:pos + 8 + k])
        leads.append((lead, trail))
        pos += 8 + k
    assert leads == [(-10, 10), (-10, -10), (3, -3)]
    with open(p, "rb") as f:
        assert R.read_record(f) == payload
        assert R.read_record(f) == b"xy"
    out = np.zeros(23, dtype=np.uint8)
    with open(p, "rb") as f:
        R._read_record_into(f, out)
    assert bytes(out) == payload


def test_dimension_mismatch_is_refused(tmp_path):
    """READ_FLOW stops on a file of another grid (src/diablo.F90:955-958)."""
    R.save_flow(str(tmp_path), _random_flow(16, 5, 7, 0))
    with pytest.raises(ValueError, match="wrong dimensions"):
        R.read_flow(str(tmp_path), 16, 5, 9)
    shutil.rmtree(tmp_path / "last_saved")
