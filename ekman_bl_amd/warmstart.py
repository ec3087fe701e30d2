"""Learned initial guess of the per-mode pressure solves (the library's
ekm_wsnet_* entry points), and the workflow around it: dump RHS/solution
pairs from solver runs, train the network on the GPU with the library's own
forward, backward and Adam kernels, attach it to a solver.

The reference has no counterpart: MUDPACK starts every solve from zero
(iguess = 0, phi = 0; src/mg_solver2d.F90:157,190, src/mud2cr.f:93,238).  With
proj_guess = 2 the consistent projection starts each real solve from the
network's guess and iterates to the same tolerance as from zero, so the
multigrid decides when the solve is done, whatever the guess; the V-cycles saved are measured by
comparing cycle counts (`eval_cycles`).  The network predicts a correction to
the previous substep's solution scaled by the ratio of the two substeps'
H_BAR (the start of proj_guess 1): a network that outputs zero is that start.

All bulk data stays in device memory (`DeviceArray`, allocated through the
library); only parameters and scalars cross to the host.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import solver as S

_vp = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
_bound = False


def _lib():
    global _bound
    L = S.lib()
    if not _bound:
        L.ekm_wsnet_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.ekm_wsnet_destroy.argtypes = [_vp]
        L.ekm_wsnet_nparams.argtypes = [_vp]
        L.ekm_wsnet_nparams.restype = ctypes.c_long
        L.ekm_wsnet_set_params.argtypes = [_vp, _F]
        L.ekm_wsnet_get_params.argtypes = [_vp, _F]
        L.ekm_wsnet_set_norm.argtypes = [_vp, ctypes.c_double, ctypes.c_double]
        L.ekm_wsnet_get_norm.argtypes = [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.ekm_wsnet_prepare.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
        L.ekm_wsnet_target.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]
        L.ekm_wsnet_forward.argtypes = [_vp, _vp, ctypes.c_int, _vp, _vp]
        L.ekm_wsnet_gradient.argtypes = [_vp, _vp, _vp, ctypes.c_int, _F, _F]
        L.ekm_wsnet_train_step.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_float, _F, _vp]
        L.ekm_wsnet_guess.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]
        L.ekm_set_warm_start_net.argtypes = [_vp, _vp]
        L.ekm_solve_planes.argtypes = [_vp] + [ctypes.POINTER(_vp)] * 4 + [ctypes.POINTER(ctypes.c_int)]
        L.ekm_dev_alloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        L.ekm_dev_free.argtypes = [_vp]
        L.ekm_copy_dev.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
        L.ekm_memcpy_h2d.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.ekm_gather_rows.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_long), ctypes.c_int, ctypes.c_size_t, _vp]
        L.ekm_memcpy_d2h.argtypes = [_vp, _vp, ctypes.c_size_t]
        _bound = True
    return L


def _check(code, what):
    if code != 0:
        raise S.EkmError(code, what)


class DeviceArray:
    """A device buffer of `n` elements of numpy dtype `dtype`."""

    def __init__(self, n: int, dtype=np.float32):
        self.n, self.dtype = int(n), np.dtype(dtype)
        self.ptr = _vp()
        _check(_lib().ekm_dev_alloc(ctypes.byref(self.ptr), max(1, self.n * self.dtype.itemsize)), "ekm_dev_alloc")

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    def at(self, offset: int) -> _vp:
        """Device address of element `offset`."""
        return _vp(self.ptr.value + int(offset) * self.dtype.itemsize)

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        _check(_lib().ekm_memcpy_h2d(self.at(offset), a.ctypes.data_as(_vp), a.nbytes), "ekm_memcpy_h2d")

    def download(self, count: int | None = None, offset: int = 0) -> np.ndarray:
        out = np.empty(self.n - offset if count is None else count, dtype=self.dtype)
        _check(_lib().ekm_memcpy_d2h(out.ctypes.data_as(_vp), self.at(offset), out.nbytes), "ekm_memcpy_d2h")
        return out

    def free(self) -> None:
        if self.ptr:
            _lib().ekm_dev_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class WarmStartNet:
    """The network on the current device for planes of nz2 x ny2 points."""

    def __init__(self, nz2: int, ny2: int, bmax: int, trainable: bool = False):
        self.nz2, self.ny2, self.bmax = nz2, ny2, bmax
        self.h1, self.w1 = (ny2 + 1) // 2, (nz2 + 1) // 2
        self.h = _vp()
        _check(_lib().ekm_wsnet_create(nz2, ny2, bmax, 1 if trainable else 0, ctypes.byref(self.h)),
               "ekm_wsnet_create")

    @property
    def nparams(self) -> int:
        return int(_lib().ekm_wsnet_nparams(self.h))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.nparams, dtype=np.float32)
        _check(_lib().ekm_wsnet_get_params(self.h, out.ctypes.data_as(_F)), "ekm_wsnet_get_params")
        return out

    def set_params(self, theta: np.ndarray) -> None:
        t = np.ascontiguousarray(theta, dtype=np.float32)
        assert t.size == self.nparams
        _check(_lib().ekm_wsnet_set_params(self.h, t.ctypes.data_as(_F)), "ekm_wsnet_set_params")

    def set_norm(self, sigma: float, kx_max: float) -> None:
        _check(_lib().ekm_wsnet_set_norm(self.h, float(sigma), float(kx_max)), "ekm_wsnet_set_norm")

    def norm(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib().ekm_wsnet_get_norm(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def prepare(self, rhs, prev, kx2, n, x1: DeviceArray, x1_off: int, scale: DeviceArray, sc_off: int,
                ratio=None):
        _check(_lib().ekm_wsnet_prepare(self.h, rhs, prev, ratio, kx2, n, x1.at(x1_off), scale.at(sc_off), None),
               "ekm_wsnet_prepare")

    def target(self, phi, scale: DeviceArray, sc_off: int, n, t0: DeviceArray, t_off: int, prev=None, ratio=None):
        _check(_lib().ekm_wsnet_target(self.h, phi, prev, ratio, scale.at(sc_off), n, t0.at(t_off), None),
               "ekm_wsnet_target")

    def forward(self, x1: DeviceArray, n: int, x1_off: int = 0) -> np.ndarray:
        y = DeviceArray(n * self.h1 * self.w1)
        _check(_lib().ekm_wsnet_forward(self.h, x1.at(x1_off), n, y.ptr, None), "ekm_wsnet_forward")
        return y.download().reshape(n, self.h1, self.w1)

    def gradient(self, x1: DeviceArray, t0: DeviceArray, n: int, x1_off: int = 0, t_off: int = 0):
        g = np.empty(self.nparams, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib().ekm_wsnet_gradient(self.h, x1.at(x1_off), t0.at(t_off), n, g.ctypes.data_as(_F),
                                         ctypes.byref(loss)), "ekm_wsnet_gradient")
        return loss.value, g

    def train_step(self, x1: DeviceArray, t0: DeviceArray, n: int, lr: float, x1_off: int = 0, t_off: int = 0):
        loss = ctypes.c_float()
        _check(_lib().ekm_wsnet_train_step(self.h, x1.at(x1_off), t0.at(t_off), n, float(lr), ctypes.byref(loss),
                                           None), "ekm_wsnet_train_step")
        return loss.value

    def attach(self, s: S.DuctSolver) -> None:
        _check(_lib().ekm_set_warm_start_net(s.h, self.h), "ekm_set_warm_start_net")

    def save(self, path: str) -> None:
        sigma, kx_max = self.norm()
        np.savez_compressed(path, theta=self.get_params(), sigma=sigma, kx_max=kx_max, nz2=self.nz2, ny2=self.ny2)

    def load(self, path: str) -> None:
        z = np.load(path)
        self.set_params(z["theta"])
        self.set_norm(float(z["sigma"]), float(z["kx_max"]))

    def close(self) -> None:
        if getattr(self, "h", None):
            _lib().ekm_wsnet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass


def detach(s: S.DuctSolver) -> None:
    _check(_lib().ekm_set_warm_start_net(s.h, None), "ekm_set_warm_start_net")


def solve_planes(s: S.DuctSolver):
    """(rhs, phi, kx2, H_BAR ratio) device addresses and the solve count of
    the last projection."""
    r, p, k, q = _vp(), _vp(), _vp(), _vp()
    n = ctypes.c_int()
    _check(_lib().ekm_solve_planes(s.h, ctypes.byref(r), ctypes.byref(p), ctypes.byref(k), ctypes.byref(q),
                                   ctypes.byref(n)), "ekm_solve_planes")
    return r, p, k, q, n.value


class PairSet:
    """Training pairs on the device: network inputs x1 (from each projection's
    RHS_TMP and the previous projection's solution of the same solve, scaled
    by the substeps' H_BAR ratio) and targets t0 (the converged solution minus
    that scaled previous solution, normalised)."""

    def __init__(self, net: WarmStartNet, capacity: int):
        self.net, self.cap, self.n = net, capacity, 0
        self.x1 = DeviceArray(capacity * net.h1 * net.w1 * 4)
        self.t0 = DeviceArray(capacity * net.nz2 * net.ny2)
        self.scale = DeviceArray(capacity, np.float64)
        self._prev = None

    def reset_run(self):
        self._prev = None

    def dump(self, s: S.DuctSolver) -> int:
        """Add the pairs of the projection the solver just did (after the
        first projection of a run, which has no previous solution)."""
        s.sync()
        rhs, phi, kx2, ratio, ns = solve_planes(s)
        plane = self.net.nz2 * self.net.ny2
        if self._prev is None:
            self._prev = DeviceArray(ns * plane, np.float64)
            _check(_lib().ekm_copy_dev(self._prev.ptr, phi, ns * plane, None), "ekm_copy_dev")
            return 0
        take = min(ns, self.cap - self.n)
        if take > 0:
            self.net.prepare(rhs, self._prev.ptr, kx2, take, self.x1, self.n * self.net.h1 * self.net.w1 * 4,
                             self.scale, self.n, ratio=ratio)
            self.net.target(phi, self.scale, self.n, take, self.t0, self.n * plane, prev=self._prev.ptr, ratio=ratio)
            self.n += take
        _check(_lib().ekm_copy_dev(self._prev.ptr, phi, ns * plane, None), "ekm_copy_dev")
        return take

    def active(self) -> np.ndarray:
        """Indices of pairs with a nonzero right-hand side."""
        return np.nonzero(self.scale.download(self.n) > 0)[0]


def run_and_dump(p, xp, yp, ic, pairs: PairSet, steps: int, warmup: int = 1) -> None:
    """One seeded run of the consistent projection from zero starts: `warmup`
    steps, then `steps` steps whose projections are dumped."""
    q = p.copy()
    q.proj_guess = 0
    s = S.DuctSolver(q, xp, yp)
    s.set_initial_physical(*ic)
    for _ in range(warmup):
        s.time_step(1)
    pairs.reset_run()
    pairs.dump(s)   # the last warm-up projection: the previous solutions
    for _ in range(steps):
        for _q in range(3):
            s.rk_duct()
            pairs.dump(s)
    s.close()


def estimate_sigma(pairs: PairSet, idx: np.ndarray, net: WarmStartNet) -> float:
    """sigma such that the normalised targets are of order one: the median
    over pairs of max|phi - ratio prev| / max|rhs| (the targets were formed
    with sigma = 1)."""
    plane = net.nz2 * net.ny2
    vals = []
    for i in idx[:: max(1, len(idx) // 256)]:
        t = pairs.t0.download(plane, int(i) * plane)
        vals.append(np.abs(t).max())
    return float(np.median(vals))


def train(net: WarmStartNet, pairs: PairSet, idx: np.ndarray, epochs: int, batch: int, lr: float,
          seed: int = 0, log=print) -> list:
    """Adam over shuffled batches of the pairs `idx` (gathered into a batch
    buffer on the device), cosine-decayed learning rate; returns the mean loss
    per epoch."""
    rng = np.random.default_rng(seed)
    h1w1, plane = net.h1 * net.w1 * 4, net.nz2 * net.ny2
    bx, bt = DeviceArray(batch * h1w1), DeviceArray(batch * plane)
    L = _lib()
    nsteps = epochs * (len(idx) // batch)
    step, hist = 0, []
    for ep in range(epochs):
        perm = rng.permutation(idx)
        losses = []
        t0 = time.perf_counter()
        for b in range(len(perm) // batch):
            sel = np.ascontiguousarray(perm[b * batch:(b + 1) * batch], dtype=np.int64)
            ip = sel.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
            _check(L.ekm_gather_rows(bx.ptr, pairs.x1.ptr, ip, batch, h1w1 * 4, None), "ekm_gather_rows")
            _check(L.ekm_gather_rows(bt.ptr, pairs.t0.ptr, ip, batch, plane * 4, None), "ekm_gather_rows")
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * step / max(1, nsteps)))
            losses.append(net.train_step(bx, bt, batch, lr_t))
            step += 1
        hist.append(float(np.mean(losses)))
        log(f"[warm start] epoch {ep}: loss {hist[-1]:.4f} ({time.perf_counter() - t0:.1f} s)")
    return hist


def eval_cycles(p, xp, yp, ic, net: WarmStartNet | None, guess: int, steps: int, warmup: int = 1):
    """V-cycles per step and ms per step of the consistent projection with the
    given start (0 zero, 1 the previous correction scaled, 2 the network), and
    the final state."""
    q = p.copy()
    q.proj_guess = guess
    s = S.DuctSolver(q, xp, yp)
    if guess == 2:
        net.attach(s)
    s.set_initial_physical(*ic)
    for _ in range(warmup):
        s.time_step(1)
    s.sync()
    cyc, worst = 0, 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        for _q in range(3):
            s.rk_duct()
            d = s.diag()
            cyc += d.mg_cycles
            worst = max(worst, d.div_max)
    s.sync()
    el = time.perf_counter() - t0
    state = s.get_state()
    s.close()
    return {"cycles_per_step": cyc / steps, "ms_per_step": el / steps * 1e3, "div_max": worst, "state": state}
