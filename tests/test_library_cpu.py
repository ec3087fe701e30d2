"""CPU checks of the product library and its host side (no GPU needed).

- libekman_bl_amd.so loads and exports every function include/*.h declares;
- the host-built tables (metrics, multigrid coefficients of every level) are
  bit-for-bit those of the oracle;
- the product path fails loudly when the library is missing (no fallback);
- the reference run-directory parser reads the shipped input.dat fixture;
- the byte contract reproduces its worked values;
- bench.py refuses more than one GPU, and finds the measured HBM traffic of
  its headline workload in the summary under profiles/.
"""
import glob
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from ekman_bl_amd import roofline
from ekman_bl_amd import solver as S
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.params import YMAX, YMIN, ZMAX, ZMIN, EkmParams, base_params, config, read_input_dat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ekm_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("library not built")
    L = S.lib()
    names = _declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert L.ekm_version().decode().startswith("ekman_bl_amd")


def test_product_path_fails_loudly_without_library(tmp_path):
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from ekman_bl_amd import solver as S\n"
        "S.LIB_PATH = %r\n"
        "S._lib = None\n"
        "try:\n"
        "    S.lib()\n"
        "except RuntimeError as e:\n"
        "    print('raised', e)\n"
        "    sys.exit(0)\n"
        "sys.exit(1)\n" % (ROOT, str(tmp_path / "missing.so")))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "not built" in r.stdout


def _nan_equal(a, b):
    return np.array_equal(np.nan_to_num(a, nan=1e300), np.nan_to_num(b, nan=1e300))


@pytest.mark.parametrize("n,bcs", [(1, (2, 2, 2, 1)), (1, (2, 2, 2, 2)), (1, (0, 0, 2, 2)), (1, (1, 1, 1, 1)),
                                   (1, (0, 0, 1, 2)), (1, (2, 2, 0, 0)), (2, (2, 2, 2, 1))])
def test_host_tables_equal_oracle(n, bcs, oracle_lib):
    O = oracle_lib
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("library not built")
    cfg = config(n)
    p = cfg.params
    for s in range(4):
        p.mudpack_bc[s] = bcs[s]
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    xp = xp + 0.05 * np.sin(0.3 * yp)       # skewed: every metric is nonzero
    o = O.Oracle(p, xp, yp)
    hm = S.host_metrics(p.nz, p.ny, xp, yp)
    for w in range(7):
        assert _nan_equal(o.metric(w), hm[2 * w:2 * w + 2]), w
    assert _nan_equal(o.metric(7), hm[14])
    lm = S.host_les_metrics(p.nz, p.ny, xp, yp)
    for w in range(6):
        assert _nan_equal(o.metric(8 + w), lm[w]), w
    for lev in range(1, o.mg_ngrid() + 1):
        for kx2 in (0.0, 3.7):
            assert np.array_equal(o.mg_discretize(kx2, lev)[:9], S.host_mg_coefficients(p, xp, yp, lev, kx2)), \
                (lev, kx2)


def test_read_input_dat_fixture():
    """rundir_unst/input.dat (fixture) through the reference's line order
    (src/diablo.F90:242-337)."""
    p = base_params(256, 127, 319, n_th=1)
    vals = read_input_dat(os.path.join(GOLDEN, "input_unst.dat"), p)
    assert abs(vals["version"] - 1.21) < 1e-12
    assert p.nu == pytest.approx(1.3e-4) and p.lx == 60 and p.ly == 60 and p.lz == 60
    assert vals["n_time_steps"] == 30000
    assert p.delta_t == 3.0 and p.variable_dt == 1 and p.cfl == pytest.approx(0.8) and p.update_dt == 1
    assert list(p.mudpack_bc) == [0, 0, 2, 2]
    assert vals["ic_type"] == 7
    assert p.f_type == 1 and p.f0 == pytest.approx(1.0e-4) and p.omega0 == pytest.approx(0.0146)
    for c in range(3):
        assert p.u_bc[c][YMIN] == 0 and p.u_bc[c][YMAX] == 1
        assert p.u_bc[c][ZMIN] == 6 and p.u_bc[c][ZMAX] == 6
    assert p.th_bc[YMIN] == 5 and p.th_bc_c1[YMIN] == pytest.approx(1.56e-3)
    assert p.th_bc[YMAX] == 1 and p.th_bc[ZMIN] == 6 and p.th_bc[ZMAX] == 6
    assert p.pr == 1.0
    assert vals["les"] is False and p.les_model == 0 and vals["wall_model_type"] == 3


def test_read_grid_def_expressions(tmp_path):
    """PARAMETER statements in the forms a grid_def uses: one or several
    names per statement, integer expressions, names defined on earlier lines,
    nested parentheses inside an expression, intrinsic calls (skipped) and
    trailing comments."""
    from ekman_bl_amd.params import params_from_run_dir, read_grid_def
    text = "\n".join([
        "      INTEGER NX, NY, NZ",
        "\tPARAMETER (NX = 64)",
        "\tPARAMETER (NY = 33-2)   ! interior rows",
        "\tPARAMETER (NZ = 65-2)",
        "\tPARAMETER (N_TH = 1)",
        "\tPARAMETER (NZ_S_R = 900)",
        "\tPARAMETER (NZ_S_L = NZ + 1 - NZ_S_R)",
        "\tPARAMETER (NY_S_T = 25)",
        "\tPARAMETER (NP= 4)",
        "\tPARAMETER (NZV=ceiling(real(NZ+2)/real(NP))*NP)",
        "\tPARAMETER (iixp = 2 , jjyq = 2, iiex = 6, jjey = 5 )",
        "\tPARAMETER (nnx=iixp*2**(iiex-1)+1, nny=jjyq*2**(jjey-1)+1)",
    ]) + "\n"
    path = tmp_path / "grid_def"
    path.write_text(text)
    g = read_grid_def(str(path))
    assert (g["nx"], g["ny"], g["nz"], g["n_th"]) == (64, 31, 63, 1)
    assert g["nz_s_l"] == 63 + 1 - 900 and g["ny_s_t"] == 25
    assert (g["iixp"], g["jjyq"], g["iiex"], g["jjey"]) == (2, 2, 6, 5)
    assert g["nnx"] == 65 and g["nny"] == 33
    assert "nzv" not in g
    p = params_from_run_dir(str(path), os.path.join(GOLDEN, "input_unst.dat"))
    assert (p.nx, p.ny, p.nz, p.ny_s_t, p.nz_s_r) == (64, 31, 63, 25, 900)
    assert (p.mg_ixp, p.mg_iex, p.mg_jyq, p.mg_jey) == (2, 6, 2, 5)
    assert list(p.mudpack_bc) == [0, 0, 2, 2]


def _edited_input(tmp_path, edit) -> str:
    lines = open(os.path.join(GOLDEN, "input_unst.dat")).read().splitlines()
    path = tmp_path / "input.dat"
    path.write_text("\n".join(edit(lines)) + "\n")
    return str(path)


def test_read_input_dat_les_line(tmp_path):
    """LES = .TRUE. with LES_MODEL_TYPE 2 (and TH 2) selects the dynamic
    Smagorinsky model; types the solver cannot honour raise, as does the
    pairing the reference stops on (src/diablo.F90:254-258)."""
    def les(model, model_th):
        def e(lines):
            i = next(n for n, ln in enumerate(lines) if ".FALSE." in ln and "2" in ln and n < 12)
            lines[i] = f"              .TRUE.        {model}               {model_th}\t\t\t3"
            return lines
        return e
    p = base_params(256, 127, 319, n_th=1)
    vals = read_input_dat(_edited_input(tmp_path, les(2, 2)), p)
    assert vals["les"] is True and p.les_model == 2
    for model, model_th in ((2, 1), (1, 1), (1, 2), (3, 3)):
        with pytest.raises(ValueError):
            read_input_dat(_edited_input(tmp_path, les(model, model_th)), base_params(256, 127, 319, n_th=1))


def test_read_input_dat_by_record_position(tmp_path):
    """A comment record is skipped whatever it holds (tab-led, 'c:', blank),
    as READ(11,*) with no items does; a blank record inside a data read holds
    no values and the read continues to the next record."""
    ref = base_params(256, 127, 319, n_th=1)
    want = read_input_dat(os.path.join(GOLDEN, "input_unst.dat"), ref)

    def retag(lines):
        out = []
        for n, ln in enumerate(lines):
            if ln.lstrip().upper().startswith("C"):
                ln = ("C\t" if n % 3 == 0 else "c:" if n % 3 == 1 else "") + ln.lstrip()[1:]
                if n % 3 == 2:
                    ln = ""           # an empty comment record
            out.append(ln)
        return out

    p = base_params(256, 127, 319, n_th=1)
    got = read_input_dat(_edited_input(tmp_path, retag), p)
    assert got == want
    assert bytes(p) == bytes(ref)

    def blank_before_values(lines):
        i = lines.index(next(ln for ln in lines if ln.strip().startswith("1.3E-4")))
        return lines[:i] + [""] + lines[i:]

    p2 = base_params(256, 127, 319, n_th=1)
    assert read_input_dat(_edited_input(tmp_path, blank_before_values), p2) == want
    assert bytes(p2) == bytes(ref)


def test_params_struct_matches_header():
    """ekm_params_default fills the C struct the ctypes mirror describes."""
    if not os.path.exists(S.LIB_PATH):
        pytest.skip("library not built")
    p = EkmParams()
    S.lib().ekm_params_default(p)
    assert p.mg_maxcy == 6 and p.mg_tolmax == 1e-6      # src/mg_solver2d.F90:158,188
    assert list(p.mudpack_bc) == [2, 2, 2, 2]
    assert p.cmplx_fp32 == 0 and p.coriolis_mean_only == 0 and p.pr == 1.0


def test_byte_contract_worked_values():
    """B_step = 8 N_pts [(34F+33) + 3 r (15 + 30.5 c)] per point-step for
    the 256x193x257 stratified case: 2,578 / 3,561 / 4,545 B at c = 2/4/6."""
    p = config(2).params
    per_pt = [roofline.step_bytes(p, c) / p.npts for c in (2, 4, 6)]
    assert [round(v) for v in per_pt] == [2578, 3561, 4545]
    p1 = config(1).params
    assert [round(roofline.step_bytes(p1, c) / p1.npts) for c in (2, 4, 6)] == [2334, 3340, 4347]


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_finds_measured_traffic_of_headline_workload():
    """The headline bench line carries roofline.traffic from the committed
    rocprofv3 summary of its configuration: each summary's workload string
    must be the one bench.py builds for that configuration's headline run."""
    sys.path.insert(0, ROOT)
    import bench
    from ekman_bl_amd.params import config
    found = 0
    for n in (2, 3, 5):
        cfg = config(n)
        if not os.path.exists(bench.traffic_file(cfg)):
            continue
        tr = bench.measured_traffic(cfg, bench.workload_name(cfg, 2, 1e-12, 1))
        assert tr is not None and tr[0] > 1e9 and tr[1] > 0, n
        found += 1
    assert found >= 1


def test_fortran_module_binds_every_entry_point():
    """Every extern "C" entry point that include/ekman_bl_amd.h declares has a
    BIND(C) interface of the same name in the Fortran module the reference's
    driver uses (ekman_bl_amd/fortran/ekman_bl_amd_mod.f90), and the module
    names no symbol the header lacks."""
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    h = open(os.path.join(root, "include", "ekman_bl_amd.h")).read()
    f = open(os.path.join(root, "ekman_bl_amd", "fortran", "ekman_bl_amd_mod.f90")).read()
    decl = set(re.findall(r"\b(ekm_\w+)\s*\(", re.sub(r"/\*.*?\*/", "", h, flags=re.S)))
    bound = set(re.findall(r'bind\(c,\s*name="(ekm_\w+)"\)', f))
    assert decl - bound == set()
    assert bound - decl == set()


def test_autotune_cache_round_trip(tmp_path):
    """The autotuner's cache keeps one entry per (workload, form, device)."""
    from ekman_bl_amd import autotune as AT
    path = str(tmp_path / "tuning.json")
    a = {"tuning": {"adi_rows": 12, "adi_block": 128}, "ms_per_step": 1.0, "trials": []}
    b = {"tuning": {"adi_rows": 4, "adi_block": 256}, "ms_per_step": 2.0, "trials": []}
    AT.save("w", 2, a, path, device="d1")
    AT.save("w", 2, b, path, device="d2")
    assert AT.load("w", 2, path, device="d1") == a["tuning"]
    assert AT.load("w", 2, path, device="d2") == b["tuning"]
    assert AT.load("w", 1, path, device="d1") is None
    assert AT.load("w", 2, str(tmp_path / "none.json"), device="d1") is None
