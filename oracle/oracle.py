"""ctypes wrapper of the CPU oracle (oracle/libekman_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle the HIP path is tested against.  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg, nowhere else.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ekman_bl_amd.params import EkmParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libekman_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libmudpack_ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_lib = None


class EkmDiag(ctypes.Structure):
    _fields_ = [("div_ref", ctypes.c_double), ("div_max", ctypes.c_double),
                ("delta_t", ctypes.c_double), ("time", ctypes.c_double),
                ("time_step", ctypes.c_int), ("rk_step", ctypes.c_int),
                ("mg_solves", ctypes.c_int), ("mg_cycles", ctypes.c_int),
                ("mg_max_cycles", ctypes.c_int), ("status", ctypes.c_int),
                ("mg_unconverged", ctypes.c_int)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.ora_create.restype = ctypes.c_void_p
        L.ora_create.argtypes = [ctypes.POINTER(EkmParams), _D, _D, _D]
        L.ora_destroy.argtypes = [ctypes.c_void_p]
        for n in ("ora_rk_duct", "ora_rem_div_curv"):
            getattr(L, n).argtypes = [ctypes.c_void_p]
        L.ora_time_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ora_set_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_int]
        L.ora_get_state.argtypes = [ctypes.c_void_p, _D, _D, _D, _D, _D]
        L.ora_get_les.argtypes = [ctypes.c_void_p, _D, _D]
        L.ora_get_les_th.argtypes = [ctypes.c_void_p, _D, _D]
        L.ora_get_forcing.argtypes = [ctypes.c_void_p, _D, _D]
        L.ora_get_sponge.argtypes = [ctypes.c_void_p, _D]
        L.ora_get_sponge_out.argtypes = [ctypes.c_void_p, _D]
        L.ora_mean_grads.argtypes = [ctypes.c_void_p, _D]
        L.ora_xmoments.argtypes = [ctypes.c_void_p, _D]
        L.ora_omega_x.argtypes = [ctypes.c_void_p, _D]
        L.ora_get_physical.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.ora_wall_model.argtypes = [ctypes.c_void_p, _D, _D, ctypes.c_int, _D, _D]
        L.ora_bc_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, _D]
        L.ora_xstats.argtypes = [ctypes.c_void_p, _D]
        L.ora_adi_field.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_double, _D, ctypes.c_double]
        L.ora_set_initial_physical.argtypes = [ctypes.c_void_p, _D, _D, _D, _D]
        L.ora_get_phi.argtypes = [ctypes.c_void_p, _D]
        L.ora_get_diag.argtypes = [ctypes.c_void_p, ctypes.POINTER(EkmDiag)]
        L.ora_get_metric.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.ora_fft_fwd.argtypes = [_D, ctypes.c_int, ctypes.c_long]
        L.ora_fft_inv.argtypes = [_D, ctypes.c_int, ctypes.c_long]
        L.ora_thomas_real.argtypes = [_D, _D, _D, _D, ctypes.c_int, ctypes.c_int]
        L.ora_thomas_real_sp.argtypes = [_D, _D, _D, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ora_mudpack.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D]
        L.ora_mg_discretize.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _D,
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ora_mg_ngrid.argtypes = [ctypes.c_void_p]
        L.ora_tight_solve.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D]
        L.ora_tight_operator.argtypes = [ctypes.c_void_p, ctypes.c_double, _D]
        L.ora_cb_bind.argtypes = [ctypes.c_void_p, ctypes.c_double]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _c(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


# -- components ---------------------------------------------------------------
def fft_fwd(varp: np.ndarray) -> None:
    """In place on (..., NX+2) float64 pencils."""
    nx = varp.shape[-1] - 2
    lib().ora_fft_fwd(_p(varp), nx, varp.size // (nx + 2))


def fft_inv(varp: np.ndarray) -> None:
    nx = varp.shape[-1] - 2
    lib().ora_fft_inv(_p(varp), nx, varp.size // (nx + 2))


def thomas_real(a, b, c, g) -> None:
    """Arrays of shape (NY+1, NX+1) (system index slow, batch fast), in place."""
    ny, nx = g.shape[0] - 1, g.shape[1] - 1
    lib().ora_thomas_real(_p(a), _p(b), _p(c), _p(g), ny, nx)


def thomas_real_sp(a, b, c, g, d: float, k: int) -> None:
    ny, nx = g.shape[0] - 1, g.shape[1] - 1
    lib().ora_thomas_real_sp(_p(a), _p(b), _p(c), _p(g), d, k, ny, nx)


# -- the duct solver ----------------------------------------------------------
class Oracle:
    def __init__(self, p: EkmParams, xpoint: np.ndarray, ypoint: np.ndarray, th_bar=None):
        self.p = p.copy()
        self._xp = np.ascontiguousarray(xpoint, dtype=np.float64)
        self._yp = np.ascontiguousarray(ypoint, dtype=np.float64)
        self._tb = None if th_bar is None else np.ascontiguousarray(th_bar, dtype=np.float64)
        h = lib().ora_create(ctypes.byref(self.p), _p(self._xp), _p(self._yp),
                             _p(self._tb) if self._tb is not None else None)
        if not h:
            raise ValueError("oracle: unsupported parameters")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def _zeros_c(self):
        return np.zeros(self.p.shape_fourier, dtype=np.complex128)

    def set_state(self, cu1, cu2, cu3, cp, cth=None, time=0.0, time_step=0):
        arrs = [np.ascontiguousarray(a, dtype=np.complex128) for a in (cu1, cu2, cu3, cp)]
        th = np.ascontiguousarray(cth if cth is not None else self._zeros_c(), dtype=np.complex128)
        st = lib().ora_set_state(self.h, *[_c(a) for a in arrs], _c(th), float(time), int(time_step))
        if st:
            raise RuntimeError(f"ora_set_state: {st}")

    def get_les(self):
        """(NU_T physical (NY+2, NZ+2, NX+2), C_DYN (NY+2, NZ+2))."""
        p = self.p
        nut = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        cdyn = np.zeros((p.ny + 2, p.nz + 2))
        st = lib().ora_get_les(self.h, _p(nut), _p(cdyn))
        if st != 0:
            raise RuntimeError(f"ora_get_les: {st}")
        return nut, cdyn

    XSTATS = ("u1", "u2", "u3", "p", "th", "urms", "vrms", "wrms", "thrms", "uv", "uw", "wv", "pu", "pv", "thv", "thw")

    def x_stats(self) -> dict:
        """x statistics by direct sums in physical space: name -> (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((16, p.ny + 2, p.nz + 2))
        lib().ora_xstats(self.h, _p(out))
        return dict(zip(self.XSTATS, out))

    MEAN_GRADS = ("dudy", "dwdy", "dudz", "dwdz", "dthdy", "dthdz", "rig", "tke")
    XMOMENTS = ("uu", "vv", "ww", "uuu", "vvv", "www", "uuuu", "vvvv", "wwww", "uuv", "wwv", "uuuv", "wwwv")

    def x_moments(self):
        """high_oder_momentums_curvi's x moments: name -> (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((13, p.ny + 2, p.nz + 2))
        lib().ora_xmoments(self.h, _p(out))
        return dict(zip(self.XMOMENTS, out))

    def omega_x(self) -> np.ndarray:
        """SAVE_STATS_CURVI's mean streamwise vorticity: (NY+2, NZ+2)."""
        p = self.p
        out = np.zeros((p.ny + 2, p.nz + 2))
        lib().ora_omega_x(self.h, _p(out))
        return out

    def physical(self, f: int) -> np.ndarray:
        """Field f (0..4: U1, U2, U3, P, TH) in physical space: (NY+2, NZ+2, NX+2)."""
        p = self.p
        out = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        lib().ora_get_physical(self.h, f, _p(out))
        return out

    def wall_model(self, u1: np.ndarray, u3: np.ndarray, first: bool = False):
        """LES_WALL_MODEL (type 3) on physical U1, U3 ((NY+2, NZ+2, NX+2)):
        (row values (2, NZ+2, NX+2) of U1 and U3, (U_1, W_1, U_STAR, TAO_WALL,
        TAO_WALL_1))."""
        p = self.p
        out = np.zeros((2, p.nz + 2, p.nx + 2))
        info = np.zeros(5)
        a, b = np.ascontiguousarray(u1, dtype=np.float64), np.ascontiguousarray(u3, dtype=np.float64)
        st = lib().ora_wall_model(self.h, _p(a), _p(b), int(first), _p(out), _p(info))
        if st != 0:
            raise ValueError("the oracle has no wall-model rows")
        return out, info

    def bc_rows(self, comp: int, direction: int, line: int, u: np.ndarray, sentinel: float = np.nan):
        """The boundary rows of one ADI line (ora_bc_rows): (a, b, c, g), each
        (rows, NX+2), the entries that are not boundary rows at `sentinel`."""
        p = self.p
        n = (p.ny + 2) if direction == 0 else (p.nz + 2)
        out = [np.full((n, p.nx + 2), sentinel) for _ in range(4)]
        uu = np.ascontiguousarray(u, dtype=np.float64)
        st = lib().ora_bc_rows(self.h, comp, direction, line, _p(uu), *(_p(a) for a in out))
        if st != 0:
            raise ValueError("ora_bc_rows failed")
        return out

    def mean_grads(self):
        """SAVE_STATS_CURVI's mean gradients: (dict name -> (NY+2, NZ+2), THBULK)."""
        p = self.p
        out = np.zeros(8 * (p.ny + 2) * (p.nz + 2) + 1)
        lib().ora_mean_grads(self.h, _p(out))
        return dict(zip(self.MEAN_GRADS, out[:-1].reshape(8, p.ny + 2, p.nz + 2))), out[-1]

    def sponge_sigma(self):
        """SPONGE_SIGMA(0:NY+1) of the vertical sponge."""
        out = np.zeros(self.p.ny + 2)
        lib().ora_get_sponge(self.h, _p(out))
        return out

    def sponge_sigma_out(self):
        """SPONGE_SIGMA_OUT(0:NZ+1) of the streamwise sponge."""
        out = np.zeros(self.p.nz + 2)
        lib().ora_get_sponge_out(self.h, _p(out))
        return out

    def forcing(self):
        """(PX0, Qn) of the F_TYPE 0 flow-rate controller."""
        a, b = ctypes.c_double(), ctypes.c_double()
        lib().ora_get_forcing(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def get_les_th(self):
        """(KAPPA_T physical (NY+2, NZ+2, NX+2), C_DYN_TH (NY+2, NZ+2))."""
        p = self.p
        kap = np.zeros((p.ny + 2, p.nz + 2, p.nx + 2))
        cdt = np.zeros((p.ny + 2, p.nz + 2))
        st = lib().ora_get_les_th(self.h, _p(kap), _p(cdt))
        if st != 0:
            raise RuntimeError(f"ora_get_les_th: {st}")
        return kap, cdt

    def get_state(self):
        out = [self._zeros_c() for _ in range(5)]
        lib().ora_get_state(self.h, *[_c(a) for a in out])
        return out  # cu1, cu2, cu3, cp, cth

    def set_initial_physical(self, u1, u2, u3, th):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u1, u2, u3, th)]
        return lib().ora_set_initial_physical(self.h, *[_p(a) for a in arrs])

    def rk_duct(self) -> int:
        return lib().ora_rk_duct(self.h)

    def time_step(self, n: int = 1) -> int:
        return lib().ora_time_step(self.h, n)

    def rem_div_curv(self) -> int:
        return lib().ora_rem_div_curv(self.h)

    def phi(self):
        out = self._zeros_c()
        lib().ora_get_phi(self.h, _c(out))
        return out

    def diag(self) -> EkmDiag:
        d = EkmDiag()
        lib().ora_get_diag(self.h, ctypes.byref(d))
        return d

    def adi_field(self, comp: int, u: np.ndarray, r: np.ndarray, t1: float, nut=None, t2: float = 0.0):
        """One field's ADI (physical (NY+2, NZ+2, NX+2) arrays, returned as new
        arrays), with the subgrid viscosity nut when given."""
        u = np.array(u, dtype=np.float64, order="C")
        r = np.array(r, dtype=np.float64, order="C")
        nu = None if nut is None else np.ascontiguousarray(nut, dtype=np.float64)
        st = lib().ora_adi_field(self.h, comp, _p(u), _p(r), float(t1), None if nu is None else _p(nu), float(t2))
        if st != 0:
            raise RuntimeError(f"ora_adi_field: {st}")
        return u, r

    def metric(self, which: int) -> np.ndarray:
        """0..6: GMAT_11, GMAT_12, GMAT_22, CJOB_11, CJOB_12, CJOB_21, CJOB_22 as
        (2, NY+3, NZ+3); 7: INT_JACOB as (NY+3, NZ+3); 8..13: the LES variants
        GMAT_11_z, GMAT_12_z, GMAT_22_z, GMAT_11_y, GMAT_12_y, GMAT_22_y."""
        n3z, n3y = self.p.nz + 3, self.p.ny + 3
        out = np.zeros((2, n3y, n3z) if which != 7 else (n3y, n3z))
        lib().ora_get_metric(self.h, which, _p(out))
        return out

    def mudpack(self, kx2: float, rhs_tmp: np.ndarray):
        """One MULTIGRID_MUDPACK solve; rhs_tmp is (NY+2, NZ+2).  Returns (phi, cycles)."""
        r = np.ascontiguousarray(rhs_tmp, dtype=np.float64)
        phi = np.zeros_like(r)
        cyc = lib().ora_mudpack(self.h, float(kx2), _p(r), _p(phi))
        return phi, cyc

    def tight_solve(self, kx2: float, rhs_tmp: np.ndarray):
        """One real solve of the consistent projection system (proj_mode 1);
        rhs_tmp is (NY+2, NZ+2).  Returns (phi, V cycles)."""
        r = np.ascontiguousarray(rhs_tmp, dtype=np.float64)
        phi = np.zeros_like(r)
        it = lib().ora_tight_solve(self.h, float(kx2), _p(r), _p(phi))
        return phi, it

    def tight_operator(self, kx2: float) -> np.ndarray:
        """Fine-level operator of proj_mode 1: (9, NY+2, NZ+2)."""
        cf = np.zeros((9, self.p.ny + 2, self.p.nz + 2))
        lib().ora_tight_operator(self.h, float(kx2), _p(cf))
        return cf

    def mg_ngrid(self) -> int:
        return lib().ora_mg_ngrid(self.h)

    def mg_discretize(self, kx2: float, lev: int) -> np.ndarray:
        nz, ny = self.p.nz + 2, self.p.ny + 2
        cf = np.zeros(nz * ny * 10)
        nx_ = ctypes.c_int()
        ny_ = ctypes.c_int()
        n = lib().ora_mg_discretize(self.h, float(kx2), lev, _p(cf), ctypes.byref(nx_), ctypes.byref(ny_))
        if n < 0:
            raise ValueError("bad level")
        return cf[: n * 10].reshape(10, ny_.value, nx_.value)


# -- reference MUDPACK (oracle/_ref) ------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_mudpack(o: Oracle, kx2: float, rhs_tmp: np.ndarray, maxcy: int = 6, tolmax: float = 1e-6):
    """Solve with the reference's own mud2cr (compiled from src/mud2cr.f and
    src/mudcom.f), driven the way MULTIGRID_MUDPACK drives it
    (src/mg_solver2d.F90:75-201), with the oracle's cofcr/bndcr callbacks."""
    R = ctypes.CDLL(REF_LIB)
    L = lib()
    p = o.p
    nnx, nny = p.nz + 2, p.ny + 2
    iprm = (ctypes.c_int * 17)()
    fprm = (ctypes.c_double * 6)()
    mgopt = (ctypes.c_int * 4)(2, 2, 1, 3)
    ierr = ctypes.c_int(0)
    iprm[0] = 0
    iprm[1], iprm[2], iprm[3], iprm[4] = p.mudpack_bc[0], p.mudpack_bc[1], p.mudpack_bc[2], p.mudpack_bc[3]
    iprm[5], iprm[6], iprm[7], iprm[8] = p.mg_ixp, p.mg_jyq, p.mg_iex, p.mg_jey
    iprm[9], iprm[10] = nnx, nny
    iprm[11] = 0         # iguess
    iprm[12] = maxcy
    iprm[13] = 0         # method
    llwork = (7 * (nnx + 2) * (nny + 2) + 44 * nnx * nny) // 3
    iprm[14] = llwork
    fprm[0], fprm[1], fprm[2], fprm[3], fprm[4] = 0.0, float(nnx - 1), 0.0, float(nny - 1), tolmax
    work = np.zeros(llwork)
    rhs = np.ascontiguousarray(-rhs_tmp, dtype=np.float64)   # phi_bc
    phi = np.zeros_like(rhs)
    L.ora_cb_bind(o.h, float(kx2))
    cof = ctypes.cast(L.ora_cofcr_cb, ctypes.c_void_p)
    bnd = ctypes.cast(L.ora_bndcr_cb, ctypes.c_void_p)
    fn = R.mud2cr_
    fn.argtypes = [ctypes.c_void_p] * 9
    args = [ctypes.cast(iprm, ctypes.c_void_p), ctypes.cast(fprm, ctypes.c_void_p),
            ctypes.c_void_p(work.ctypes.data), cof, bnd, ctypes.c_void_p(rhs.ctypes.data),
            ctypes.c_void_p(phi.ctypes.data), ctypes.cast(mgopt, ctypes.c_void_p),
            ctypes.cast(ctypes.byref(ierr), ctypes.c_void_p)]
    fn(*args)
    e0 = ierr.value
    iprm[0] = 1
    fn(*args)
    return phi, int(iprm[16]), (e0, ierr.value)
