"""Run parameters of the duct time step.

`EkmParams` mirrors ``ekm_params`` in include/ekman_bl_amd.h field for field.
It holds the part of the reference's compile-time ``grid_def`` and runtime
``input.dat`` that the RK3 substep reads (reference: src/diablo.F90:242-337,
rundir_unst/grid_def:6-44).  `read_input_dat` / `read_grid_def` parse those
two files with the reference's fixed line order, so a run directory of the
reference can be pointed at this package unchanged.

`config(n)` builds the five benchmark configurations (neutral 64x65x65,
stratified 256x193x257, 512x257x513 and 1024x385x769) on the reference's
boundary-condition codes.
"""
from __future__ import annotations

import ctypes
import math
import re
from dataclasses import dataclass

YMIN, YMAX, ZMIN, ZMAX = 0, 1, 2, 3


class EkmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("n_th", ctypes.c_int),
        ("mg_ixp", ctypes.c_int), ("mg_jyq", ctypes.c_int),
        ("mg_iex", ctypes.c_int), ("mg_jey", ctypes.c_int),
        ("nu", ctypes.c_double), ("lx", ctypes.c_double),
        ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("delta_t", ctypes.c_double),
        ("variable_dt", ctypes.c_int),
        ("cfl", ctypes.c_double),
        ("update_dt", ctypes.c_int),
        ("f_type", ctypes.c_int),
        ("wbulk0", ctypes.c_double), ("px0", ctypes.c_double),
        ("omega0", ctypes.c_double), ("amp_omega0", ctypes.c_double),
        ("f0", ctypes.c_double),
        ("u_bc", (ctypes.c_int * 4) * 3),
        ("u_bc_c1", (ctypes.c_double * 4) * 3),
        ("ri_tau", ctypes.c_double), ("pr", ctypes.c_double),
        ("cont_strat", ctypes.c_int),
        ("th_bc", ctypes.c_int * 4),
        ("th_bc_c1", ctypes.c_double * 4),
        ("mudpack_bc", ctypes.c_int * 4),
        ("mg_maxcy", ctypes.c_int),
        ("mg_tolmax", ctypes.c_double),
        ("cmplx_fp32", ctypes.c_int),
        ("coriolis_mean_only", ctypes.c_int),
        ("proj_mode", ctypes.c_int),
        ("proj_maxit", ctypes.c_int),
        ("proj_tol", ctypes.c_double),
        ("proj_guess", ctypes.c_int),
        ("proj_krylov", ctypes.c_int),
        ("les_model", ctypes.c_int),
        ("sponge_type", ctypes.c_int),
        ("ny_s_t", ctypes.c_int),
        ("nz_s_r", ctypes.c_int),
        ("save_stats_int", ctypes.c_int),
        ("nz_s_l", ctypes.c_int),
        ("wall_model_type", ctypes.c_int),
    ]

    # -- convenience ---------------------------------------------------------
    @property
    def shape_phys(self):
        """(NY+2, NZ+2, NX+2): numpy shape of a physical field (x fastest)."""
        return (self.ny + 2, self.nz + 2, self.nx + 2)

    @property
    def shape_fourier(self):
        """(NY+2, NZ+2, NX/2+1): numpy shape of a complex Fourier field."""
        return (self.ny + 2, self.nz + 2, self.nx // 2 + 1)

    @property
    def npts(self):
        """Grid points as the metric counts them: NX * (NY+2) * (NZ+2)."""
        return self.nx * (self.ny + 2) * (self.nz + 2)

    def copy(self) -> "EkmParams":
        p = EkmParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self), ctypes.sizeof(EkmParams))
        return p


def mg_factors(n: int) -> tuple[int, int]:
    """(ixp, iex) with n = ixp * 2**(iex-1) + 1 and the smallest ixp >= 2
    (MUDPACK's size rule, src/mud2cr.f:148-150)."""
    m = n - 1
    iex = 1
    while m % 2 == 0 and m // 2 >= 2:
        m //= 2
        iex += 1
    if m < 2:
        raise ValueError(f"{n} points do not fit MUDPACK's ixp*2**(iex-1)+1 rule")
    return m, iex


def base_params(nx: int, ny: int, nz: int, n_th: int = 0) -> EkmParams:
    p = EkmParams()
    p.nx, p.ny, p.nz, p.n_th = nx, ny, nz, n_th
    p.mg_ixp, p.mg_iex = mg_factors(nz + 2)
    p.mg_jyq, p.mg_jey = mg_factors(ny + 2)
    p.nu, p.lx, p.ly, p.lz = 1.0e-3, 4 * math.pi, 20.0, 2 * math.pi
    p.delta_t, p.variable_dt, p.cfl, p.update_dt = 0.1, 1, 0.5, 1
    p.f_type = 3          # no mean-flow pressure forcing
    p.pr = 1.0
    p.mg_maxcy, p.mg_tolmax = 6, 1.0e-6
    for s in range(4):
        p.mudpack_bc[s] = 2
    p.proj_maxit, p.proj_tol, p.proj_krylov = 50, 1.0e-12, 1   # used by proj_mode 1, 2
    p.ny_s_t = p.nz_s_r = p.save_stats_int = 1 << 30             # no sponge rows
    p.nz_s_l = -(1 << 30)
    p.wall_model_type = 3                                        # rundir_unst/input.dat:11
    return p


@dataclass
class Config:
    name: str
    params: EkmParams
    gamma: float          # tanh stretching factor of the y grid
    steps: int            # step count the configuration is quoted on
    seed: int
    noise: float          # uniform noise amplitude, fraction of G
    th_noise: float


def _ekman_bcs(p: EkmParams, ug: float) -> None:
    """Neutral/stratified Ekman layer on the reference's BC codes (IC_TYPE 7):
    no-slip bottom, geostrophic/free-slip top, free-slip z ends."""
    U1, U2, U3 = 0, 1, 2
    p.u_bc[U1][YMIN] = 0; p.u_bc_c1[U1][YMIN] = 0.0
    p.u_bc[U2][YMIN] = 0; p.u_bc_c1[U2][YMIN] = 0.0
    p.u_bc[U3][YMIN] = 0; p.u_bc_c1[U3][YMIN] = 0.0
    # top: U1 carries the geostrophic wind (its C1 doubles as the Coriolis
    # offset, src/duct.F90:471), V = 0, W stress-free
    p.u_bc[U1][YMAX] = 0; p.u_bc_c1[U1][YMAX] = ug
    p.u_bc[U2][YMAX] = 0; p.u_bc_c1[U2][YMAX] = 0.0
    p.u_bc[U3][YMAX] = 1; p.u_bc_c1[U3][YMAX] = 0.0
    # free-slip z ends: U1, V Neumann, W = 0
    for s in (ZMIN, ZMAX):
        p.u_bc[U1][s] = 1; p.u_bc[U2][s] = 1; p.u_bc[U3][s] = 0
    # pressure: Neumann at the wall and the z ends (z ends use the generalised
    # bndcr), Dirichlet at the top so that the kx = 0 problem is not singular
    p.mudpack_bc[0] = 2
    p.mudpack_bc[1] = 2
    p.mudpack_bc[2] = 2
    p.mudpack_bc[3] = 1
    # geostrophic offsets on the mean mode only (the reference subtracts them
    # from every kx mode, src/duct.F90:471-473; both forms are supported)
    p.coriolis_mean_only = 1


def config(n: int) -> Config:
    """The benchmark configurations 1-5 (4: configuration 2 with the learned initial guess)."""
    if n == 1:
        p = base_params(64, 63, 63, n_th=0)
        re_d = 500.0
        p.nu = 1.0 / re_d
        p.f0 = 2.0 / re_d
        p.lx, p.ly, p.lz = 4 * math.pi, 20.0, 2 * math.pi
        _ekman_bcs(p, 1.0)
        p.delta_t, p.cfl = 0.2, 0.5
        return Config("config1_neutral_64x65x65", p, 2.0, 100, 1, 0.05, 0.0)
    if n == 4:
        # the grid and flow of configuration 2 with the learned initial guess
        # of every consistent pressure solve; its network is trained on
        # pairs from seeds 0-9 and it runs held-out seed 10
        c2 = config(2)
        p = c2.params.copy()
        p.proj_mode, p.proj_guess = 2, 2
        return Config("config4_learned_256x193x257", p, c2.gamma, c2.steps, 10, c2.noise, c2.th_noise)
    if n in (2, 3, 5):
        dims = {2: (256, 191, 255), 3: (512, 255, 511), 5: (1024, 383, 767)}[n]
        re_d = {2: 1000.0, 3: 1000.0, 5: 5000.0}[n]
        nf = {2: 30.0, 3: 50.0, 5: 50.0}[n]
        p = base_params(*dims, n_th=1)
        p.nu = 1.0 / re_d
        p.f0 = 2.0 / re_d
        p.lx, p.ly, p.lz = 4 * math.pi, 20.0, 2 * math.pi
        _ekman_bcs(p, 1.0)
        p.pr = 0.7
        p.cont_strat = 1
        p.ri_tau = (nf * p.f0) ** 2           # N^2 = RI_TAU with TH_BAR = -y
        p.th_bc[YMIN] = 1; p.th_bc_c1[YMIN] = -1.0e-3   # surface cooling flux
        p.th_bc[YMAX] = 1; p.th_bc_c1[YMAX] = 0.0
        p.th_bc[ZMIN] = 1; p.th_bc[ZMAX] = 1
        p.delta_t, p.cfl = 0.1, 0.5
        if n == 5:
            p.les_model = 2   # dynamic Smagorinsky for the velocity and the scalar
        steps = {2: 1000, 3: 500, 5: 2000}[n]
        return Config(f"config{n}_stratified_{dims[0]}x{dims[1]+2}x{dims[2]+2}", p, 2.0, steps, 2, 0.05, 0.01)
    raise ValueError(f"no configuration {n}")


# ---------------------------------------------------------------------------
# reference input files
# ---------------------------------------------------------------------------
def read_grid_def(path: str) -> dict:
    """PARAMETER values of a grid_def file (NX, NY, NZ, N_TH, iixp, ...).
    Every ``name = expression`` of a PARAMETER statement whose expression,
    after the names defined before it are replaced by their values, holds
    only integers and + - * / ** ( ) is evaluated (``129-2``,
    ``NZ + 1 - NZ_S_R``, ``iixp*2**(iiex-1)+1``); the others (intrinsic
    calls such as ``ceiling(real(NZ+2)/real(NP))*NP``) are skipped."""
    out = {}
    for line in open(path):
        line = line.split("!")[0]
        m = re.search(r"PARAMETER\s*\((.*)\)", line, re.I)
        if not m:
            continue
        # split the statement at its top-level commas
        parts, depth, cur = [], 0, ""
        for ch in m.group(1):
            depth += (ch == "(") - (ch == ")")
            if ch == "," and depth == 0:
                parts.append(cur)
                cur = ""
            else:
                cur += ch
        parts.append(cur)
        for part in parts:
            if "=" not in part:
                continue
            name, expr = (x.strip() for x in part.split("=", 1))
            expr = re.sub(r"[A-Za-z_]\w*", lambda t: str(out.get(t.group(0).lower(), t.group(0))), expr)
            if re.fullmatch(r"[0-9+\-*/ ()]+", expr):
                try:
                    out[name.lower()] = int(eval(expr))  # noqa: S307 - digits and operators only
                except (SyntaxError, ZeroDivisionError):
                    pass
    return out


def _fortran_tokens(line: str) -> list[str]:
    """Values of one record of list-directed input: separators are blanks,
    tabs and commas; a slash ends the record's values."""
    return line.split("/")[0].replace(",", " ").split()


def _fval(tok: str):
    t = tok.strip().lower()
    if t.startswith(".t") or t in ("t", "true"):
        return True
    if t.startswith(".f") or t in ("f", "false"):
        return False
    t = t.replace("d", "e")
    if re.fullmatch(r"[+-]?\d+", t):
        return int(t)
    return float(t)


class _Records:
    """The record cursor of a Fortran list-directed READ(11,*) sequence:
    `skip()` is READ(11,*) with no items (one record, whatever it holds);
    `read(n)` is READ(11,*) of n items, which takes values from as many
    records as it needs (a blank record holds none) and drops the rest of
    its last record."""

    def __init__(self, lines: list[str], path: str):
        self.lines, self.pos, self.path = lines, 0, path

    def _next(self) -> str:
        if self.pos >= len(self.lines):
            raise ValueError(f"{self.path}: end of file after record {self.pos}")
        self.pos += 1
        return self.lines[self.pos - 1]

    def skip(self) -> None:
        self._next()

    def read(self, n: int) -> list:
        vals: list = []
        while len(vals) < n:
            rec = self.pos + 1
            try:
                vals += [_fval(t) for t in _fortran_tokens(self._next())]
            except ValueError as e:
                raise ValueError(f"{self.path}: record {rec}: {e}") from None
        return vals[:n]


def read_input_dat(path: str, p: EkmParams) -> dict:
    """Fill `p` from an input.dat (version 1.21) by the reference's record
    positions (src/diablo.F90:242-337): four header records, then each value
    line preceded by one comment record, whatever that record holds, and the
    10 records per scalar when N_TH = 1.  Returns every value read.

    The LES line sets `p.les_model`: LES = .TRUE. with LES_MODEL_TYPE 2 is the
    dynamic Smagorinsky model (and, with a scalar, LES_MODEL_TYPE_TH must be 2
    as well); other model types raise ValueError, as does the combination the
    reference stops on (src/diablo.F90:256-258)."""
    r = _Records(open(path).read().splitlines(), path)
    out: dict = {}
    for _ in range(4):
        r.skip()
    out["version"] = r.read(1)[0]
    if abs(float(out["version"]) - 1.21) > 1e-6:
        raise ValueError(f"{path}: input file version {out['version']}, expected 1.21 (src/diablo.F90:248)")
    r.skip()
    p.nu, p.lx, p.ly, p.lz = r.read(4)
    r.skip()
    nts, dt, reset, vdt, cfl, dfn, upd = r.read(7)
    p.delta_t, p.variable_dt, p.cfl, p.update_dt = dt, int(vdt), cfl, int(upd)
    out.update(n_time_steps=nts, reset_time=reset, dfn=dfn)
    r.skip()
    les, les_type, les_type_th, wall_type = r.read(4)
    out.update(les=les, les_model_type=les_type, les_model_type_th=les_type_th, wall_model_type=wall_type)
    if les and les_type == 1 and les_type_th == 2:
        raise ValueError(f"{path}: LES_MODEL_TYPE=1 with LES_MODEL_TYPE_TH=2 (the reference stops, "
                         "src/diablo.F90:256-258)")
    r.skip()
    mud_use, a, b, c, d = r.read(5)
    out["mudpack_use"] = mud_use
    p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3] = a, b, c, d
    r.skip()
    io = r.read(5)  # VERBOSITY, SAVE_FLOW_INT, SAVE_STATS_INT, BACKUP_FLOW_INT, MOVIE
    p.save_stats_int = int(io[2])
    out.update(save_flow_int=io[1], save_stats_int=io[2], backup_flow_int=io[3])
    r.skip()
    ics = r.read(5)  # CREATE_NEW_FLOW, IC_TYPE, KICK, INT_TREAT, WRITE_VEL_TECPLOT
    out.update(create_new_flow=ics[0], ic_type=ics[1], kick=ics[2])
    r.skip()
    ft, wb, px0, om, amp, _beta, f0, sponge = r.read(8)
    p.f_type, p.wbulk0, p.px0, p.omega0, p.amp_omega0, p.f0 = ft, wb, px0, om, amp, f0
    out["sponge_type"] = sponge
    p.sponge_type = int(sponge)
    # BCs: U,V,W lower; U,V,W upper; U,V,W left; U,V,W right (:284-320)
    for slot in (YMIN, YMAX, ZMIN, ZMAX):
        for comp in range(3):
            r.skip()
            row = r.read(4)
            p.u_bc[comp][slot] = row[0]
            p.u_bc_c1[comp][slot] = row[1]
    r.skip()
    out["stochastic_forcing"] = r.read(1)[0]
    r.skip()
    out["filter_vel"] = r.read(2)
    r.skip()
    if p.n_th:
        r.skip()
        out["create_new_th"] = r.read(1)[0]
        r.skip()
        out["filter_th"] = r.read(2)
        r.skip()
        ri, pr, bg, cs = r.read(4)
        p.ri_tau, p.pr, p.cont_strat = ri, pr, int(cs)
        out["background_grad"] = bg
        for slot in (YMIN, YMAX, ZMIN, ZMAX):
            r.skip()
            row = r.read(4)
            p.th_bc[slot] = row[0]
            p.th_bc_c1[slot] = row[1]
    if les:
        if les_type != 2:
            raise ValueError(f"{path}: LES_MODEL_TYPE {les_type} is not supported (dynamic Smagorinsky, "
                             "type 2, only)")
        if p.n_th and les_type_th != 2:
            raise ValueError(f"{path}: LES_MODEL_TYPE_TH {les_type_th} is not supported with a scalar "
                             "(dynamic model, type 2, only)")
        p.les_model = 2
    else:
        p.les_model = 0
    p.wall_model_type = int(wall_type)
    return out


def params_from_fixture(grid_json: str, input_dat: str):
    """(params, xpoint, ypoint) of a run directory kept as a fixture pair: a
    JSON file with grid_def's PARAMETER values under "grid_def" and the
    GRID_YZ.dat coordinates as the tanh family's end points ("y", "z";
    gamma 2.3, the shipped grid's family), and the run's input.dat."""
    import json

    from .grid import tanh_grid
    d = json.load(open(grid_json))
    g = d["grid_def"]
    p = base_params(g["nx"], g["ny"], g["nz"], n_th=g.get("n_th", 0))
    read_input_dat(input_dat, p)
    p.mg_ixp, p.mg_iex, p.mg_jyq, p.mg_jey = g["iixp"], g["iiex"], g["jjyq"], g["jjey"]
    for key in ("ny_s_t", "nz_s_r", "nz_s_l"):
        if key in g:
            setattr(p, key, g[key])
    xp, yp = tanh_grid(p.ny, p.nz, d["y"][-1], d["z"][-1], 2.3)
    return p, xp, yp


def params_from_run_dir(grid_def: str, input_dat: str) -> EkmParams:
    g = read_grid_def(grid_def)
    p = base_params(g["nx"], g["ny"], g["nz"], n_th=g.get("n_th", 0))
    if "iixp" in g:
        p.mg_ixp, p.mg_jyq, p.mg_iex, p.mg_jey = g["iixp"], g["jjyq"], g["iiex"], g["jjey"]
    for key in ("ny_s_t", "nz_s_r", "nz_s_l"):
        if key in g:
            setattr(p, key, g[key])
    read_input_dat(input_dat, p)
    return p
