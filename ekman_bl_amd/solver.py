"""Python binding of libekman_bl_amd.so (the gfx950 duct step).

`DuctSolver` mirrors the reference's drop-in surface: `rk_duct()` is one
RK_DUCT substep (src/duct.F90:237), `time_step(n)` runs n steps of three
substeps (src/diablo.F90:152-157), `rem_div_curv()` is the projection
(src/duct.F90:2472), and the module-level functions wrap the component entry
points FFT_X_TO_FOURIER_OP / FFT_X_TO_PHYSICAL_OP, THOMAS_REAL(_SP) and
MULTIGRID_MUDPACK.  Status codes replace the reference's STOP statements and
raise `EkmError` here.

The library is loaded from this package directory; a missing library raises
at first use: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import EkmParams

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libekman_bl_amd.so")
_D = ctypes.POINTER(ctypes.c_double)
_lib = None

ERRORS = {-1: "bad argument", -2: "HIP runtime error", -3: "unsupported option",
          -4: "DIV is a NaN", -5: "DELTA_T is zero", -6: "no HIP device"}


class EkmError(RuntimeError):
    def __init__(self, code: int, what: str):
        msg = f"{what}: {ERRORS.get(code, code)} ({code})"
        if code == -2 and _lib is not None:
            msg += ": " + _lib.ekm_last_error().decode()
        super().__init__(msg)
        self.code = code


class EkmDiag(ctypes.Structure):
    _fields_ = [("div_ref", ctypes.c_double), ("div_max", ctypes.c_double),
                ("delta_t", ctypes.c_double), ("time", ctypes.c_double),
                ("time_step", ctypes.c_int), ("rk_step", ctypes.c_int),
                ("mg_solves", ctypes.c_int), ("mg_cycles", ctypes.c_int),
                ("mg_max_cycles", ctypes.c_int), ("status", ctypes.c_int),
                ("mg_unconverged", ctypes.c_int)]


class EkmTuning(ctypes.Structure):
    """ekm_tuning: launch and memory-placement parameters (bit-identical results)."""
    _fields_ = [("adi_rows", ctypes.c_int), ("adi_block", ctypes.c_int), ("adi_les_rows", ctypes.c_int),
                ("mg_block", ctypes.c_int), ("tight_lds_kib", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    @classmethod
    def from_dict(cls, d: dict) -> "EkmTuning":
        t = default_tuning()
        for k, v in d.items():
            setattr(t, k, int(v))
        return t


def default_tuning() -> EkmTuning:
    t = EkmTuning()
    lib().ekm_tuning_default(ctypes.byref(t))
    return t


def lib():
    """Load libekman_bl_amd.so; raises when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m ekman_bl_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.ekm_version.restype = ctypes.c_char_p
        L.ekm_last_error.restype = ctypes.c_char_p
        L.ekm_params_default.argtypes = [ctypes.POINTER(EkmParams)]
        L.ekm_create.argtypes = [ctypes.POINTER(EkmParams), _D, _D, _D, ctypes.POINTER(vp)]
        L.ekm_create_tuned.argtypes = [ctypes.POINTER(EkmParams), _D, _D, _D, ctypes.POINTER(EkmTuning),
                                       ctypes.POINTER(vp)]
        L.ekm_tuning_default.argtypes = [ctypes.POINTER(EkmTuning)]
        L.ekm_get_tuning.argtypes = [vp, ctypes.POINTER(EkmTuning)]
        L.ekm_destroy.argtypes = [vp]
        L.ekm_sync.argtypes = [vp]
        L.ekm_rk_duct.argtypes = [vp]
        L.ekm_rk_duct_graph_launch.argtypes = [vp]
        L.ekm_time_step_ex.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.ekm_rem_div_curv.argtypes = [vp]
        L.ekm_time_step.argtypes = [vp, ctypes.c_int]
        L.ekm_set_state.argtypes = [vp, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int]
        L.ekm_get_state.argtypes = [vp, _D, _D, _D, _D, _D]
        L.ekm_get_les.argtypes = [vp, _D, _D]
        L.ekm_get_les_th.argtypes = [vp, _D, _D]
        L.ekm_get_solve_cycles.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.ekm_get_forcing.argtypes = [vp, _D, _D]
        L.ekm_get_xstats.argtypes = [vp, _D]
        L.ekm_get_mean_grads.argtypes = [vp, _D]
        L.ekm_get_xmoments.argtypes = [vp, _D]
        L.ekm_get_omega_x.argtypes = [vp, _D]
        L.ekm_get_plane.argtypes = [vp, ctypes.c_int, ctypes.c_int, _D]
        L.ekm_get_physical.argtypes = [vp, ctypes.c_int, _D]
        L.ekm_set_initial_physical.argtypes = [vp, _D, _D, _D, _D]
        L.ekm_get_phi.argtypes = [vp, _D]
        L.ekm_get_diag.argtypes = [vp, ctypes.POINTER(EkmDiag)]
        L.ekm_field_device_ptr.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
        L.ekm_multigrid_mudpack.argtypes = [vp, ctypes.c_int, _D, _D, _D, ctypes.POINTER(ctypes.c_int)]
        L.ekm_fft_x_to_fourier.argtypes = [_D, ctypes.c_int, ctypes.c_int]
        L.ekm_fft_x_to_physical.argtypes = [_D, ctypes.c_int, ctypes.c_int]
        L.ekm_fft_x_to_fourier_dev.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.ekm_fft_x_to_physical_dev.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.ekm_thomas_real.argtypes = [_D, _D, _D, _D, ctypes.c_int, ctypes.c_int]
        L.ekm_thomas_real_sp.argtypes = [_D, _D, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ekm_thomas_real_dev.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp]
        L.ekm_host_metrics.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]
        L.ekm_host_les_metrics.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D, _D]
        L.ekm_host_sponge_sigma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]
        L.ekm_host_sponge_sigma_out.argtypes = [ctypes.c_int] * 5 + [_D, _D]
        L.ekm_host_mg_coefficients.argtypes = [ctypes.POINTER(EkmParams), _D, _D, ctypes.c_int, ctypes.c_double,
                                               _D, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _check(code: int, what: str) -> None:
    if code != 0:
        raise EkmError(code, what)


def _p(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_D)


def _c(a: np.ndarray):
    if a.dtype != np.complex128 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("expected a C-contiguous complex128 array")
    return a.ctypes.data_as(_D)


class DuctSolver:
    """Device-resident duct state on the current HIP device."""

    def __init__(self, p: EkmParams, xpoint: np.ndarray, ypoint: np.ndarray, th_bar=None, tuning=None):
        self.p = p.copy()
        xp = np.ascontiguousarray(xpoint, dtype=np.float64)
        yp = np.ascontiguousarray(ypoint, dtype=np.float64)
        tb = None if th_bar is None else np.ascontiguousarray(th_bar, dtype=np.float64)
        h = ctypes.c_void_p()
        if isinstance(tuning, dict):
            tuning = EkmTuning.from_dict(tuning)
        _check(lib().ekm_create_tuned(ctypes.byref(self.p), _p(xp), _p(yp), _p(tb) if tb is not None else None,
                                      ctypes.byref(tuning) if tuning is not None else None, ctypes.byref(h)),
               "ekm_create")
        self.h = h

    def tuning(self) -> EkmTuning:
        t = EkmTuning()
        _check(lib().ekm_get_tuning(self.h, ctypes.byref(t)), "ekm_get_tuning")
        return t

    def close(self):
        if getattr(self, "h", None):
            lib().ekm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def _zc(self):
        return np.zeros(self.p.shape_fourier, dtype=np.complex128)

    def set_state(self, cu1, cu2, cu3, cp, cth=None, time=0.0, time_step=0):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (cu1, cu2, cu3, cp)]
        th = np.ascontiguousarray(cth if cth is not None else self._zc(), dtype=np.complex128)
        _check(lib().ekm_set_state(self.h, *[_c(a) for a in arrs], _c(th), float(time), int(time_step)),
               "ekm_set_state")

    def get_state(self):
        out = [self._zc() for _ in range(5)]
        _check(lib().ekm_get_state(self.h, *[_c(a) for a in out]), "ekm_get_state")
        return out  # cu1, cu2, cu3, cp, cth

    def save_flow(self, run_dir: str, kind: str = "saved", backup_index=None):
        """Write the state as the reference's SAVE_FLOW / BACKUP_FLOW would
        (src/diablo.F90:753-916) under run_dir; returns the paths."""
        from . import restart
        cu1, cu2, cu3, cp, cth = self.get_state()
        d = self.diag()
        p = self.p
        flow = restart.Flow(p.nx, p.ny, p.nz, d.time, d.time_step, cu1, cu2, cu3, cp, [cth] if p.n_th else [])
        return restart.save_flow(run_dir, flow, kind, backup_index)

    def read_flow(self, run_dir: str, nranks: int = 1, kind: str = "saved", backup_index=None):
        """Restart from the files READ_FLOW reads (src/diablo.F90:918-1000),
        written by nranks ranks; a restart in the sense of set_state."""
        from . import restart
        p = self.p
        fl = restart.read_flow(run_dir, p.nx, p.ny, p.nz, p.n_th, nranks, kind, backup_index)
        self.set_state(fl.cu1, fl.cu2, fl.cu3, fl.cp, fl.cth[0] if p.n_th else None, time=fl.time,
                       time_step=fl.time_step)
        return fl

    def get_les(self):
        """(NU_T physical (NY+2, NZ+2, NX+2), C_DYN (NY+2, NZ+2)) of the subgrid model."""
        p = self.p
        nut = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        cdyn = np.zeros((p.ny + 2, p.nz + 2))
        _check(lib().ekm_get_les(self.h, _p(nut), _p(cdyn)), "ekm_get_les")
        return nut, cdyn

    XSTATS = ("u1", "u2", "u3", "p", "th", "urms", "vrms", "wrms", "thrms", "uv", "uw", "wv", "pu", "pv", "thv", "thw")

    def x_stats(self) -> dict:
        """x statistics of the current state (ekm_get_xstats): name -> (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((16, p.ny + 2, p.nz + 2))
        _check(lib().ekm_get_xstats(self.h, _p(out)), "ekm_get_xstats")
        return dict(zip(self.XSTATS, out))

    MEAN_GRADS = ("dudy", "dwdy", "dudz", "dwdz", "dthdy", "dthdz", "rig", "tke")
    XMOMENTS = ("uu", "vv", "ww", "uuu", "vvv", "www", "uuuu", "vvvv", "wwww", "uuv", "wwv", "uuuv", "wwwv")

    def x_moments(self) -> dict:
        """The higher-order x moments (ekm_get_xmoments): name -> (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((13, p.ny + 2, p.nz + 2))
        _check(lib().ekm_get_xmoments(self.h, _p(out)), "ekm_get_xmoments")
        return dict(zip(self.XMOMENTS, out))

    def omega_x(self) -> np.ndarray:
        """SAVE_STATS_CURVI's mean streamwise vorticity (ekm_get_omega_x): (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((p.ny + 2, p.nz + 2))
        _check(lib().ekm_get_omega_x(self.h, _p(out)), "ekm_get_omega_x")
        return out

    PLANE_FIELDS = ("u1", "u2", "u3", "p", "th")

    def plane(self, axis: int, index: int) -> np.ndarray:
        """Physical U1, U2, U3, P, TH on one plane (ekm_get_plane): axis 0 the
        x-y plane at K = index, (5, NY+2, NX); axis 1 the z-y plane at
        I = index, (5, NY+2, NZ+2); axis 2 the x-z plane at J = index,
        (5, NZ+2, NX)."""
        p = self.p
        shape = {0: (5, p.ny + 2, p.nx), 1: (5, p.ny + 2, p.nz + 2), 2: (5, p.nz + 2, p.nx)}[axis]
        out = np.zeros(shape)
        _check(lib().ekm_get_plane(self.h, axis, index, _p(out)), "ekm_get_plane")
        return out

    def physical(self, f: int) -> np.ndarray:
        """Field f (0..4: U1, U2, U3, P, TH) in physical space (ekm_get_physical):
        (NY+2, NZ+2, NX+2), the last two x entries undefined."""
        p = self.p
        out = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        _check(lib().ekm_get_physical(self.h, f, _p(out)), "ekm_get_physical")
        return out

    def mean_grads(self):
        """SAVE_STATS_CURVI's mean gradients of the current state
        (ekm_get_mean_grads): (dict name -> (NY+2, NZ+2), THBULK)."""
        p = self.p
        out = np.zeros(8 * (p.ny + 2) * (p.nz + 2) + 1)
        _check(lib().ekm_get_mean_grads(self.h, _p(out)), "ekm_get_mean_grads")
        return dict(zip(self.MEAN_GRADS, out[:-1].reshape(8, p.ny + 2, p.nz + 2))), out[-1]

    def forcing(self):
        """(PX0, Qn) of the F_TYPE 0 flow-rate controller."""
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().ekm_get_forcing(self.h, ctypes.byref(a), ctypes.byref(b)), "ekm_get_forcing")
        return a.value, b.value

    def solve_cycles(self) -> np.ndarray:
        """Cycles of every real solve of the last projection (2 per kx mode)."""
        n = 2 * (self.p.nx // 3 + 1)
        out = np.zeros(n, dtype=np.int32)
        _check(lib().ekm_get_solve_cycles(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), n),
               "ekm_get_solve_cycles")
        return out

    def get_les_th(self):
        """(KAPPA_T physical (NY+2, NZ+2, NX+2), C_DYN_TH (NY+2, NZ+2)) of the scalar model."""
        p = self.p
        kap = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        cdt = np.zeros((p.ny + 2, p.nz + 2))
        _check(lib().ekm_get_les_th(self.h, _p(kap), _p(cdt)), "ekm_get_les_th")
        return kap, cdt

    def set_initial_physical(self, u1, u2, u3, th):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u1, u2, u3, th)]
        _check(lib().ekm_set_initial_physical(self.h, *[_p(a) for a in arrs]), "ekm_set_initial_physical")

    def rk_duct(self, graph: bool = False) -> None:
        """One substep, launched eagerly or replayed from a captured HIP graph.
        Its projection checks are read back by diag(), sync() and time_step()."""
        if graph:
            _check(lib().ekm_rk_duct_graph_launch(self.h), "RK_DUCT (graph)")
        else:
            _check(lib().ekm_rk_duct(self.h), "RK_DUCT")

    def time_step(self, n: int = 1, graph: bool = False) -> None:
        _check(lib().ekm_time_step_ex(self.h, int(n), 1 if graph else 0), "time_step")

    def rem_div_curv(self) -> None:
        _check(lib().ekm_rem_div_curv(self.h), "REM_DIV_CURV")

    def phi(self) -> np.ndarray:
        out = self._zc()
        _check(lib().ekm_get_phi(self.h, _c(out)), "ekm_get_phi")
        return out

    def sync(self) -> None:
        _check(lib().ekm_sync(self.h), "ekm_sync")

    def diag(self) -> EkmDiag:
        d = EkmDiag()
        _check(lib().ekm_get_diag(self.h, ctypes.byref(d)), "ekm_get_diag")
        return d

    def multigrid_mudpack(self, kx2: np.ndarray, rhs_tmp: np.ndarray):
        """Batched MULTIGRID_MUDPACK: rhs_tmp (nsolve, NY+2, NZ+2) -> (phi, cycles)."""
        k = np.ascontiguousarray(kx2, dtype=np.float64)
        r = np.ascontiguousarray(rhs_tmp, dtype=np.float64)
        phi = np.zeros_like(r)
        cyc = (ctypes.c_int * len(k))()
        _check(lib().ekm_multigrid_mudpack(self.h, len(k), _p(k), _p(r), _p(phi), cyc), "MULTIGRID_MUDPACK")
        return phi, np.array(cyc[:], dtype=np.int32)


# -- component drop-ins (host arrays) -------------------------------------------
def fft_x_to_fourier(varp: np.ndarray) -> None:
    """FFT_X_TO_FOURIER_OP in place on (..., NX+2) float64 pencils."""
    nx = varp.shape[-1] - 2
    _check(lib().ekm_fft_x_to_fourier(_p(varp), nx, varp.size // (nx + 2)), "FFT_X_TO_FOURIER_OP")


def fft_x_to_physical(varp: np.ndarray) -> None:
    nx = varp.shape[-1] - 2
    _check(lib().ekm_fft_x_to_physical(_p(varp), nx, varp.size // (nx + 2)), "FFT_X_TO_PHYSICAL_OP")


def thomas_real(a, b, c, g) -> None:
    """THOMAS_REAL on arrays of shape (NY+1, NX+1): solution returned in g."""
    ny, nx = g.shape[0] - 1, g.shape[1] - 1
    _check(lib().ekm_thomas_real(_p(a), _p(b), _p(c), _p(g), ny, nx), "THOMAS_REAL")


def thomas_real_sp(a, b, c, g, d: float, k: int) -> None:
    ny, nx = g.shape[0] - 1, g.shape[1] - 1
    _check(lib().ekm_thomas_real_sp(_p(a), _p(b), _p(c), _p(g), float(d), int(k), ny, nx), "THOMAS_REAL_SP")


def host_metrics(nz: int, ny: int, xpoint: np.ndarray, ypoint: np.ndarray) -> np.ndarray:
    """Metric tables as the library computes them (no device needed):
    (15, NY+3, NZ+3) = GMAT_11,12,22, CJOB_11,12,21,22 (two faces each) + INT_JACOB."""
    mn = (nz + 3) * (ny + 3)
    out = np.zeros(15 * mn)
    _check(lib().ekm_host_metrics(nz, ny, _p(np.ascontiguousarray(xpoint, dtype=np.float64)),
                                  _p(np.ascontiguousarray(ypoint, dtype=np.float64)), _p(out)), "ekm_host_metrics")
    return out.reshape(15, ny + 3, nz + 3)


def host_sponge_sigma(sponge_type: int, nz: int, ny: int, ny_s_t: int, ypoint: np.ndarray) -> np.ndarray:
    """SPONGE_SIGMA(0:NY+1) of the vertical sponge (src/diablo.F90:1533-1556)."""
    out = np.zeros(ny + 2)
    _check(lib().ekm_host_sponge_sigma(sponge_type, nz, ny, ny_s_t,
                                       _p(np.ascontiguousarray(ypoint, dtype=np.float64)), _p(out)),
           "ekm_host_sponge_sigma")
    return out


def host_sponge_sigma_out(sponge_type: int, nz: int, ny: int, nz_s_r: int, nz_s_l: int,
                          xpoint: np.ndarray) -> np.ndarray:
    """SPONGE_SIGMA_OUT(0:NZ+1) of the streamwise sponge (src/diablo.F90:1489-1523)."""
    out = np.zeros(nz + 2)
    _check(lib().ekm_host_sponge_sigma_out(sponge_type, nz, ny, nz_s_r, nz_s_l,
                                           _p(np.ascontiguousarray(xpoint, dtype=np.float64)), _p(out)),
           "ekm_host_sponge_sigma_out")
    return out


def host_les_metrics(nz: int, ny: int, xpoint: np.ndarray, ypoint: np.ndarray) -> np.ndarray:
    """Subgrid-stress metric variants as the library computes them:
    (6, 2, NY+3, NZ+3) = GMAT_11_z, 12_z, 22_z, 11_y, 12_y, 22_y."""
    mn = (nz + 3) * (ny + 3)
    out = np.zeros(12 * mn)
    _check(lib().ekm_host_les_metrics(nz, ny, _p(np.ascontiguousarray(xpoint, dtype=np.float64)),
                                      _p(np.ascontiguousarray(ypoint, dtype=np.float64)), _p(out)),
           "ekm_host_les_metrics")
    return out.reshape(6, 2, ny + 3, nz + 3)


def host_mg_coefficients(p: EkmParams, xpoint, ypoint, lev: int, kx2: float) -> np.ndarray:
    """cf(i,j,1..9) of multigrid level `lev` as the library builds it: (9, ny, nx)."""
    cf = np.zeros((p.nz + 2) * (p.ny + 2) * 9)
    nx = ctypes.c_int()
    ny = ctypes.c_int()
    _check(lib().ekm_host_mg_coefficients(ctypes.byref(p), _p(np.ascontiguousarray(xpoint, dtype=np.float64)),
                                          _p(np.ascontiguousarray(ypoint, dtype=np.float64)), lev, float(kx2),
                                          _p(cf), ctypes.byref(nx), ctypes.byref(ny)), "ekm_host_mg_coefficients")
    n = nx.value * ny.value
    return cf[: 9 * n].reshape(9, ny.value, nx.value)


# -- measurement hooks ------------------------------------------------------------
def prof_enable(s: DuctSolver, on: bool = True) -> None:
    L = lib()
    L.ekm_prof_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    _check(L.ekm_prof_enable(s.h, 1 if on else 0), "ekm_prof_enable")


def prof_get(s: DuctSolver):
    """(ms, launches, solves, cycles) of the multigrid launches timed since prof_enable."""
    L = lib()
    L.ekm_prof_get.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)] + [ctypes.POINTER(ctypes.c_long)] * 3
    ms = ctypes.c_double()
    a, b, c = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
    _check(L.ekm_prof_get(s.h, ctypes.byref(ms), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)), "ekm_prof_get")
    return ms.value, a.value, b.value, c.value
