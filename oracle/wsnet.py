"""CPU restatement of the learned initial guess of the per-mode Poisson solves
(the library's ekm_wsnet, ekman_bl_amd/csrc/ekm_wsnet.hip), in float64 numpy.

TEST INFRASTRUCTURE ONLY: the oracle the HIP network kernels are checked
against.  Imported by tests/ only.

The reference has no counterpart: its MUDPACK solves always start from zero
(iguess = 0, phi = 0; src/mg_solver2d.F90:157,190; the iguess branch of
src/mud2cr.f:93,238 is never taken).  The network predicts that start.

Layout: activations are (batch, H, W, C) with the plane's K (z) index as W
(fastest) and J (y) as H.  A 3x3 convolution with stride S and zero padding
maps (H, W) to ((H-1)//S + 1, (W-1)//S + 1):

    y[n, oy, ox, co] = b[co] + sum_{ci, dy, dx} W[co, ci, t] x[n, S oy + dy, S ox + dx, ci],
    t = 3 (dy + 1) + (dx + 1),  dy, dx in {-1, 0, 1}.

With `bf16=True` the operands the GPU feeds to its bf16 matrix cores are
rounded to bfloat16 (round to nearest even) as the GPU rounds them: the input
activation and the weight in the forward product, the output gradient and
the weight in the input-gradient product, the output gradient and the input
activation in the weight-gradient product.  Everything else (bias, ReLU,
1x1 output layer, upsampling, loss) is exact float64.
"""
from __future__ import annotations

import numpy as np

# (kind, stride, inputs, cout, relu, output): the U-Net of the library, in
# execution order.  Tensors: 'x' input (4 channels, level 1); s0..s3 the
# encoder outputs kept for the skips; b the bottleneck; u3..u0 upsampled
# decoder tensors; y the prediction at level 1.
CH = (8, 16, 32, 48, 64)
NET = [
    ("conv", 1, ("x",), CH[0], True, "e0"),
    ("conv", 1, ("e0",), CH[0], True, "s0"),
    ("conv", 2, ("s0",), CH[1], True, "e1"),
    ("conv", 1, ("e1",), CH[1], True, "s1"),
    ("conv", 2, ("s1",), CH[2], True, "e2"),
    ("conv", 1, ("e2",), CH[2], True, "s2"),
    ("conv", 2, ("s2",), CH[3], True, "e3"),
    ("conv", 1, ("e3",), CH[3], True, "s3"),
    ("conv", 2, ("s3",), CH[4], True, "e4"),
    ("conv", 1, ("e4",), CH[4], True, "b"),
    ("up", 0, ("b",), CH[4], False, "u3"),
    ("conv", 1, ("s3", "u3"), CH[3], True, "d3a"),
    ("conv", 1, ("d3a",), CH[3], True, "d3"),
    ("up", 0, ("d3",), CH[3], False, "u2"),
    ("conv", 1, ("s2", "u2"), CH[2], True, "d2a"),
    ("conv", 1, ("d2a",), CH[2], True, "d2"),
    ("up", 0, ("d2",), CH[2], False, "u1"),
    ("conv", 1, ("s1", "u1"), CH[1], True, "d1a"),
    ("conv", 1, ("d1a",), CH[1], True, "d1"),
    ("up", 0, ("d1",), CH[1], False, "u0"),
    ("conv", 1, ("s0", "u0"), CH[0], True, "d0a"),
    ("conv", 1, ("d0a",), CH[0], True, "d0"),
    ("conv1", 1, ("d0",), 1, False, "y"),
]
CIN0 = 4


def channels() -> dict:
    c = {"x": CIN0}
    for kind, _s, ins, cout, _r, out in NET:
        c[out] = cout if kind != "up" else c[ins[0]]
    return c


def param_layout():
    """[(layer index, cout, cin, taps)] of the trainable layers in the order of
    the flat parameter vector (per layer: W[cout, cin, taps], then b[cout])."""
    c = channels()
    out = []
    for i, (kind, _s, ins, cout, _r, _o) in enumerate(NET):
        if kind == "up":
            continue
        cin = sum(c[t] for t in ins)
        out.append((i, cout, cin, 9 if kind == "conv" else 1))
    return out


def nparams() -> int:
    return sum(co * ci * t + co for _i, co, ci, t in param_layout())


def unpack(theta: np.ndarray) -> dict:
    """layer index -> (W, b) views of the flat parameter vector."""
    out, off = {}, 0
    for i, co, ci, t in param_layout():
        w = theta[off:off + co * ci * t].reshape(co, ci, t)
        off += co * ci * t
        b = theta[off:off + co]
        off += co
        out[i] = (w, b)
    assert off == theta.size
    return out


def he_init(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    parts = []
    for _i, co, ci, t in param_layout():
        parts.append(rng.standard_normal(co * ci * t) * np.sqrt(2.0 / (ci * t)))
        parts.append(np.zeros(co))
    return np.concatenate(parts)


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) through float32, as the GPU's
    conversion of its float32 operands does."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _taps():
    return [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def conv3(x, w, b, s, rnd):
    n, h, wd, _c = x.shape
    ho, wo = (h - 1) // s + 1, (wd - 1) // s + 1
    xp = np.pad(rnd(x), ((0, 0), (1, 1), (1, 1), (0, 0)))
    wr = rnd(w)
    y = np.broadcast_to(b, (n, ho, wo, w.shape[0])).copy()
    for t, (dy, dx) in enumerate(_taps()):
        xs = xp[:, 1 + dy: 1 + dy + s * (ho - 1) + 1: s, 1 + dx: 1 + dx + s * (wo - 1) + 1: s, :]
        y += xs @ wr[:, :, t].T
    return y


def conv3_bwd(x, w, gy, s, rnd):
    """(gx, gW, gb) of conv3 for the output gradient gy."""
    n, h, wd, _c = x.shape
    ho, wo = gy.shape[1], gy.shape[2]
    xp = np.pad(rnd(x), ((0, 0), (1, 1), (1, 1), (0, 0)))
    gyr, wr = rnd(gy), rnd(w)
    gxp = np.zeros((n, h + 2, wd + 2, x.shape[3]))
    gw = np.zeros_like(w)
    for t, (dy, dx) in enumerate(_taps()):
        sl = (slice(None), slice(1 + dy, 1 + dy + s * (ho - 1) + 1, s), slice(1 + dx, 1 + dx + s * (wo - 1) + 1, s))
        gw[:, :, t] = np.einsum("nhwo,nhwi->oi", gyr, xp[sl])
        gxp[sl] += gyr @ wr[:, :, t]
    return gxp[:, 1:-1, 1:-1, :], gw, gy.sum(axis=(0, 1, 2))


def up(c, shape):
    """Bilinear prolongation from (h, w) to (2h-1, 2w-1): even points copy,
    odd points average their two coarse neighbours (per direction)."""
    h, w = c.shape[1], c.shape[2]
    assert shape == (2 * h - 1, 2 * w - 1), (shape, c.shape)
    f = np.zeros((c.shape[0], 2 * h - 1, w, c.shape[3]))
    f[:, 0::2] = c
    f[:, 1::2] = 0.5 * (c[:, :-1] + c[:, 1:])
    g = np.zeros((c.shape[0], 2 * h - 1, 2 * w - 1, c.shape[3]))
    g[:, :, 0::2] = f
    g[:, :, 1::2] = 0.5 * (f[:, :, :-1] + f[:, :, 1:])
    return g


def up_adj(g):
    """Adjoint of `up`."""
    f = g[:, :, 0::2].copy()
    f[:, :, :-1] += 0.5 * g[:, :, 1::2]
    f[:, :, 1:] += 0.5 * g[:, :, 1::2]
    c = f[:, 0::2].copy()
    c[:, :-1] += 0.5 * f[:, 1::2]
    c[:, 1:] += 0.5 * f[:, 1::2]
    return c


def forward(theta, x, bf16_ops=False):
    """Prediction at level 1, (n, H1, W1, 1), and the activations."""
    rnd = bf16 if bf16_ops else (lambda a: a)
    P = unpack(theta)
    act = {"x": x}
    for i, (kind, s, ins, _cout, relu, out) in enumerate(NET):
        if kind == "up":
            src = act[ins[0]]
            skip = {"u3": "s3", "u2": "s2", "u1": "s1", "u0": "s0"}[out]
            act[out] = up(src, act[skip].shape[1:3])
            continue
        xin = np.concatenate([act[t] for t in ins], axis=3)
        w, b = P[i]
        if kind == "conv":
            y = conv3(xin, w, b, s, rnd)
        else:
            y = xin @ w[:, :, 0].T + b
        act[out] = np.maximum(y, 0.0) if relu else y
    return act["y"], act


def loss_and_grad(theta, x, t0, bf16_ops=False):
    """Mean over samples of sum (up(y) - t)^2 / sum t^2 at level 0, and its
    gradient with respect to theta."""
    rnd = bf16 if bf16_ops else (lambda a: a)
    y, act = forward(theta, x, bf16_ops)
    yf = up(y, t0.shape[1:3])
    den = (t0 ** 2).sum(axis=(1, 2, 3))
    num = ((yf - t0) ** 2).sum(axis=(1, 2, 3))
    nb = x.shape[0]
    loss = float(np.mean(num / den))
    g = {"y": up_adj(2.0 * (yf - t0) / den[:, None, None, None] / nb)}
    P = unpack(theta)
    grad = np.zeros_like(theta)
    G = unpack(grad)
    c = channels()
    for i in range(len(NET) - 1, -1, -1):
        kind, s, ins, _cout, relu, out = NET[i]
        gy = g.pop(out)
        if kind == "up":
            gsrc = up_adj(gy)
            g[ins[0]] = g.get(ins[0], 0.0) + gsrc * (act[ins[0]] > 0)
            continue
        xin = np.concatenate([act[t] for t in ins], axis=3)
        w, _b = P[i]
        if kind == "conv":
            gx, gw, gb = conv3_bwd(xin, w, gy, s, rnd)
        else:
            gw = np.einsum("nhwo,nhwi->oi", gy, xin)[:, :, None]
            gb = gy.sum(axis=(0, 1, 2))
            gx = gy @ w[:, :, 0]
        G[i][0][...] = gw
        G[i][1][...] = gb
        if ins == ("x",):
            continue
        c0 = 0
        for tname in ins:
            part = gx[..., c0:c0 + c[tname]]
            c0 += c[tname]
            # every tensor a conv reads is a ReLU output, except the upsampled ones
            mask = (act[tname] > 0) if not tname.startswith("u") else 1.0
            g[tname] = g.get(tname, 0.0) + part * mask
    return loss, grad


# -- input preparation and the guess ------------------------------------------
def restrict(v):
    """Full weighting of (n, H, W) planes to ((H+1)//2, (W+1)//2), mirrored ends."""
    def r1(a, axis):
        a = np.moveaxis(a, axis, -1)
        p = np.concatenate([a[..., 1:2], a, a[..., -2:-1]], axis=-1)
        out = 0.25 * p[..., 0:-2:2] + 0.5 * p[..., 1:-1:2] + 0.25 * p[..., 2::2]
        return np.moveaxis(out, -1, axis)
    return r1(r1(v, 1), 2)


def prepare(rhs, prev, kx, kx_max, sigma, ratio=1.0):
    """Network input (n, H1, W1, 4) and the per-solve scale s = max|rhs| from
    planes (n, NY+2, NZ+2): rhs/s, ratio prev/(s sigma), kx/kx_max and the
    row coordinate J/(H1-1) at level 1."""
    s = np.abs(rhs).max(axis=(1, 2))
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    r = restrict(rhs) * inv[:, None, None]
    q = ratio * restrict(prev) * (inv / sigma)[:, None, None]
    n, h1, w1 = r.shape
    x = np.zeros((n, h1, w1, 4))
    x[..., 0], x[..., 1] = r, q
    x[..., 2] = (np.where(s > 0, kx / kx_max, 0.0))[:, None, None]
    x[..., 3] = np.where(s[:, None, None] > 0, (np.arange(h1) / (h1 - 1))[None, :, None], 0.0)
    return x, s



def target(phi, prev, ratio, s, sigma):
    """Training target (n, NY+2, NZ+2): (phi - ratio prev) / (s sigma), zero
    where s = 0."""
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    return (phi - ratio * prev) * (inv / sigma)[:, None, None]


def guess(theta, rhs, prev, kx, kx_max, sigma, ratio=1.0, bf16_ops=False):
    """The guess of the solves: ratio prev + s sigma up(y)."""
    x, s = prepare(rhs, prev, kx, kx_max, sigma, ratio)
    y, _ = forward(theta, x, bf16_ops)
    return ratio * prev + (s * sigma)[:, None, None] * up(y, rhs.shape[1:])[..., 0]