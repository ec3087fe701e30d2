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

reference_ic7 gives the form of the reference's own IC_TYPE 7 start
(CREATE_FLOW_DUCT / CREATE_TH_DUCT): uniform W = WBULK0 with damped Fourier
kicks, scalar zero, with the same counter-based noise in place of
RANDOM_NUMBER.
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


def index_ranges(p):
    """(JSTART, JEND, ZSTART, ZEND) of INIT_DUCT for the velocity (src/duct.F90:
    27-129, IC_TYPE 7 forms): from U3's y boundary types and U1's z types."""
    wy0, wy1 = p.u_bc[2][0], p.u_bc[2][1]
    uz0, uz1 = p.u_bc[0][2], p.u_bc[0][3]
    js = 1 if wy0 in (0, 1, 3, 6, 7) else 2
    je = p.ny if wy1 in (0, 1, 6) else p.ny - 1
    zs = 1 if uz0 in (1, 6) else 2
    ze = p.nz if uz1 in (0, 1, 6) else p.nz - 1
    return js, je, zs, ze


def reference_ic7_fourier(p, kick: float, seed: int):
    """The form of CREATE_FLOW_DUCT / CREATE_TH_DUCT for IC_TYPE 7 (src/duct.F90:
    3026-3137, 2854-2864) in Fourier space, (NY+2, NZ+2, NX/2+1) complex:
    U1 = U2 = 0, U3 = WBULK0 (the mean mode), TH = 0; then for every mode
    I = 1..NX/2 and interior J, K, uniform kicks (RNUM - 1/2) KICK added to the
    real part of each velocity component (a REAL added to a COMPLEX), and the
    three components multiplied by DAMP_FACT = exp(-6 J^2 / NY^2).  The kicks
    come from the splitmix64 stream (seed, component), visited in the
    reference's I, J, K order, in place of RANDOM_NUMBER."""
    ny2, nz2, nxc = p.ny + 2, p.nz + 2, p.nx // 2 + 1
    cu = [np.zeros((ny2, nz2, nxc), dtype=np.complex128) for _ in range(3)]
    cu[2][..., 0] = p.wbulk0
    js, je, zs, ze = index_ranges(p)
    nj, nk, ni = je - js + 1, ze - zs + 1, nxc - 1
    n = ni * nj * nk
    if n > 0 and kick != 0.0:
        j = np.arange(js, je + 1)
        damp = np.exp(-6.0 * (j.astype(np.float64) ** 2) / float(p.ny ** 2))
        for c in range(3):
            r = uniform01(seed, 10 + c, n).reshape(ni, nj, nk) - 0.5   # I outer, then J, then K
            block = cu[c][js:je + 1, zs:ze + 1, 1:]
            block.real += np.transpose(r, (1, 2, 0)) * kick
            block *= damp[:, None, None]
    return cu[0], cu[1], cu[2], np.zeros_like(cu[0]), np.zeros_like(cu[0])


def fourier_to_physical(p, cf: np.ndarray) -> np.ndarray:
    """Inverse x transform in the library's convention (forward scaled 1/NX,
    modes above NX/3 dropped), padded to NX+2: (NY+2, NZ+2, NX+2)."""
    c = cf.copy()
    c[..., p.nx // 3 + 1:] = 0.0
    u = np.fft.irfft(c, n=p.nx, axis=-1) * p.nx
    out = np.zeros(u.shape[:-1] + (p.nx + 2,))
    out[..., :p.nx] = u
    return out


def reference_ic7(p, kick: float, seed: int):
    """Physical (u1, u2, u3, th) of reference_ic7_fourier, for
    set_initial_physical (which then projects, as CREATE_FLOW_DUCT's
    REM_DIV_CURV does).  Modes above NX/3 carry kicks in the reference until
    its first forward transform drops them; here they are dropped at once."""
    f = reference_ic7_fourier(p, kick, seed)
    return tuple(fourier_to_physical(p, a) for a in (f[0], f[1], f[2], f[4]))
