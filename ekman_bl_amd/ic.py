"""Seeded synthetic initial conditions.

The reference draws its noise with the compiler's RANDOM_NUMBER after
RANDOM_SEED(put=10) (src/diablo.F90:214-217, src/duct.F90:3117-3137), which
no other compiler reproduces.  Initial fields here come from a counter-based
splitmix64 stream keyed by (seed, field, linear index), so the same fields
can be regenerated anywhere and handed to both the GPU path and the oracle.

Velocity: the laminar Ekman spiral for a geostrophic wind G along x with the
reference's Coriolis sign (du/dt = -f (w - Wg), dw/dt = f (u - Ug),
src/duct.F90:471-473): u = G (1 - e^-eta cos eta), w = -G e^-eta sin eta,
eta = y / delta, delta = sqrt(2 nu / f); plus uniform noise of amplitude
`noise` * G in [-1/2, 1/2).  Scalar: zero perturbation plus `th_noise` noise
(the background TH_BAR carries the stratification).
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser of a uint64 counter array (Steele et al. 2014)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0,1) from counter (seed, stream, index)."""
    idx = np.arange(n, dtype=np.uint64)
    key = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        r = splitmix64(idx ^ splitmix64(np.full(n, key, dtype=np.uint64)))
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def ekman_initial(p, ypoint: np.ndarray, seed: int, noise: float, th_noise: float, ug: float = 1.0):
    """(u1, u2, u3, th) physical fields of shape (NY+2, NZ+2, NX+2)."""
    ny2, nz2, nx2 = p.ny + 2, p.nz + 2, p.nx + 2
    delta = np.sqrt(2.0 * p.nu / p.f0) if p.f0 > 0 else 1.0
    eta = ypoint / delta                         # (NY+2, NZ+2)
    u_base = ug * (1.0 - np.exp(-eta) * np.cos(eta))
    w_base = -ug * np.exp(-eta) * np.sin(eta)
    shape = (ny2, nz2, nx2)
    n = ny2 * nz2 * nx2
    fields = []
    for comp, base in enumerate((u_base, np.zeros_like(u_base), w_base)):
        r = uniform01(seed, comp, n).reshape(shape) - 0.5
        f = base[:, :, None] + noise * ug * r
        fields.append(f)
    th = th_noise * (uniform01(seed, 3, n).reshape(shape) - 0.5)
    u1, u2, u3 = fields
    # the padding entries I = NX, NX+1 carry no data in the reference layout
    for f in (u1, u2, u3, th):
        f[:, :, p.nx:] = 0.0
    return u1, u2, u3, th
