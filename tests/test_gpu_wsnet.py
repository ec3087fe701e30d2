"""The learned initial guess on the GPU (ekm_wsnet_*, bf16 matrix-core
kernels) against the oracle's float64 restatement (oracle/wsnet.py), and in
the solver.

- forward output and parameter gradients: the oracle rounds the matrix-core
  operands to bf16 as the GPU does, so the two differ only by fp32
  accumulation (relative L2 <= 2e-3 per layer);
- input preparation (full weighting, per-solve scale, channels): FP64
  arithmetic rounded once to FP32, within 2.5e-7 relative of the oracle;
- training with the library's own backward and Adam kernels lowers the loss
  on pairs dumped from the 64x65x65 case;
- the consistent projection started from the trained network's guess
  converges to the solution it reaches from zero (the same tolerance), and
  the cycles per step of both starts are reported.
"""
import numpy as np
import pytest

from ekman_bl_amd import solver as S
from ekman_bl_amd import warmstart as WS
from ekman_bl_amd.grid import tanh_grid
from ekman_bl_amd.ic import ekman_initial
from ekman_bl_amd.params import config
from oracle import wsnet as OW

pytestmark = pytest.mark.gpu

NZ2 = NY2 = 65   # level 1: 33 x 33, level 5: 3 x 3


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _case(seed, n=3):
    rng = np.random.default_rng(seed)
    theta = OW.he_init(seed).astype(np.float32)
    # biases nonzero so that their path is exercised
    P = OW.unpack(theta.astype(np.float64))
    th = theta.astype(np.float64)
    off = 0
    for _i, co, ci, t in OW.param_layout():
        off += co * ci * t
        th[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    theta = th.astype(np.float32)
    h1, w1 = (NY2 + 1) // 2, (NZ2 + 1) // 2
    x = rng.standard_normal((n, h1, w1, OW.CIN0)).astype(np.float32)
    t0 = rng.standard_normal((n, NY2, NZ2, 1)).astype(np.float32)
    del P
    return theta, x, t0


def test_forward_matches_oracle():
    theta, x, _ = _case(11)
    net = WS.WarmStartNet(NZ2, NY2, 4, trainable=False)
    net.set_params(theta)
    xd = WS.DeviceArray(x.size)
    xd.upload(x)
    y = net.forward(xd, x.shape[0])
    yo, _ = OW.forward(theta.astype(np.float64), x.astype(np.float64), bf16_ops=True)
    yx, _ = OW.forward(theta.astype(np.float64), x.astype(np.float64), bf16_ops=False)
    r, rx = _rel(y, yo[..., 0]), _rel(y, yx[..., 0])
    print(f"[wsnet] forward: vs bf16-operand oracle {r:.2e}, vs plain float64 {rx:.2e}", flush=True)
    assert r < 2e-3
    net.close()


def test_gradient_matches_oracle():
    theta, x, t0 = _case(12)
    net = WS.WarmStartNet(NZ2, NY2, 4, trainable=True)
    net.set_params(theta)
    xd, td = WS.DeviceArray(x.size), WS.DeviceArray(t0.size)
    xd.upload(x)
    td.upload(t0)
    loss, g = net.gradient(xd, td, x.shape[0])
    lo, go = OW.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), t0.astype(np.float64), bf16_ops=True)
    print(f"[wsnet] loss {loss:.6f} vs oracle {lo:.6f}", flush=True)
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    off, worst = 0, 0.0
    for i, co, ci, t in OW.param_layout():
        n = co * ci * t
        rw = _rel(g[off:off + n], go[off:off + n])
        rb = _rel(g[off + n:off + n + co], go[off + n:off + n + co])
        worst = max(worst, rw, rb)
        print(f"[wsnet] layer {i}: weight gradient {rw:.2e}, bias gradient {rb:.2e}", flush=True)
        off += n + co
    assert worst < 2e-3
    net.close()


def test_prepare_matches_oracle():
    rng = np.random.default_rng(13)
    n = 3
    rhs = rng.standard_normal((n, NY2, NZ2))
    rhs[1] *= 1e-3
    prev = rng.standard_normal((n, NY2, NZ2))
    kx2 = np.array([0.25, 1.0, 2.25])
    net = WS.WarmStartNet(NZ2, NY2, 4, trainable=False)
    net.set_norm(7.0, 2.0)
    r, p, k = (WS.DeviceArray(a.size, np.float64) for a in (rhs, prev, kx2))
    r.upload(rhs)
    p.upload(prev)
    k.upload(kx2)
    q = WS.DeviceArray(1, np.float64)
    q.upload(np.array([0.375]))
    x1 = WS.DeviceArray(n * 33 * 33 * 4)
    sc = WS.DeviceArray(n, np.float64)
    net.prepare(r.ptr, p.ptr, k.ptr, n, x1, 0, sc, 0, ratio=q.ptr)
    xo, so = OW.prepare(rhs, prev, np.sqrt(kx2), 2.0, 7.0, ratio=0.375)
    assert np.array_equal(sc.download(), so)
    xg = x1.download().reshape(n, 33, 33, 4)
    # the kernel forms each input in FP64 and rounds it once to FP32: within
    # two FP32 rounding units of the oracle's FP64 value (the second covers a
    # different FP64 summation order in the restriction)
    err = np.abs(xg - xo) / np.maximum(np.abs(xo), 1e-30)
    print(f"[wsnet] prepare: max relative {err.max():.2e}", flush=True)
    assert err.max() < 2.5e-7
    net.close()


def test_guess_matches_oracle():
    """ekm_wsnet_guess: ratio * prev plus s sigma times the upsampled output,
    against the oracle's guess with bf16 operands (the correction to 2e-3
    relative, the ratio * prev part exactly), written in place over prev."""
    theta, _, _ = _case(17)
    rng = np.random.default_rng(19)
    n = 3
    rhs = rng.standard_normal((n, NY2, NZ2))
    prev = rng.standard_normal((n, NY2, NZ2))
    kx2 = np.array([0.0, 1.0, 4.0])
    net = WS.WarmStartNet(NZ2, NY2, 4, trainable=False)
    net.set_params(theta)
    net.set_norm(3.0, 2.0)
    r, p, k = (WS.DeviceArray(a.size, np.float64) for a in (rhs, prev, kx2))
    r.upload(rhs)
    p.upload(prev)
    k.upload(kx2)
    q = WS.DeviceArray(1, np.float64)
    q.upload(np.array([0.625]))
    L = S.lib()
    assert L.ekm_wsnet_guess(net.h, r.ptr, p.ptr, q.ptr, k.ptr, n, p.ptr, None) == 0
    g = p.download().reshape(n, NY2, NZ2)
    go = OW.guess(theta.astype(np.float64), rhs, prev, np.sqrt(kx2), 2.0, 3.0, ratio=0.625, bf16_ops=True)
    e = _rel(g - 0.625 * prev, go - 0.625 * prev)
    print(f"[wsnet] guess: correction {e:.2e}", flush=True)
    assert e < 2e-3
    net.close()


@pytest.fixture(scope="module")
def trained():
    """A network trained for 40 epochs on the pairs of four steps of the
    64x65x65 case (seed 1), and its loss history."""
    cfg = config(1)
    p = cfg.params.copy()
    p.proj_mode = 2
    xp, yp = tanh_grid(p.ny, p.nz, p.ly, p.lz, cfg.gamma)
    ic = ekman_initial(p, yp, cfg.seed, cfg.noise, cfg.th_noise)
    nsolve = 2 * (p.nx // 3 + 1)
    net = WS.WarmStartNet(p.nz + 2, p.ny + 2, 16, trainable=True)
    kx_max = 2 * np.pi * (p.nx // 3) / p.lx
    net.set_norm(1.0, kx_max)
    pairs = WS.PairSet(net, 12 * nsolve)
    WS.run_and_dump(p, xp, yp, ic, pairs, steps=4, warmup=1)
    idx = pairs.active()
    sigma = WS.estimate_sigma(pairs, idx, net)
    # the pairs were formed with sigma = 1: dump again with the estimate
    net.set_norm(sigma, kx_max)
    pairs = WS.PairSet(net, 12 * nsolve)
    WS.run_and_dump(p, xp, yp, ic, pairs, steps=4, warmup=1)
    idx = pairs.active()
    print(f"[wsnet] {len(idx)} pairs, sigma {sigma:.4e}", flush=True)
    hist = WS.train(net, pairs, idx, epochs=40, batch=16, lr=2e-3, log=lambda m: print(m, flush=True))
    yield p, xp, yp, ic, net, hist
    net.close()


def test_training_lowers_the_loss(trained):
    _p, _xp, _yp, _ic, _net, hist = trained
    print(f"[wsnet] loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    # the targets are what the network adds to the scaled previous solution;
    # how much of it a network of this size can predict is measured by the
    # cycle counts, this only checks that the library's training descends
    assert hist[-1] < 0.97 * hist[0]


def test_learned_start_reaches_the_same_solution(trained):
    p, xp, yp, ic, net, _hist = trained
    inf = WS.WarmStartNet(p.nz + 2, p.ny + 2, 2 * (p.nx // 3 + 1), trainable=False)
    inf.set_params(net.get_params())
    inf.set_norm(*net.norm())
    ic2 = ekman_initial(p, yp, 7, 0.05, 0.0)   # another seed than the training run
    r0 = WS.eval_cycles(p, xp, yp, ic2, None, 0, steps=3)
    r2 = WS.eval_cycles(p, xp, yp, ic2, inf, 2, steps=3)
    rel = [_rel(a, b) for a, b in zip(r2["state"][:4], r0["state"][:4])]
    print(f"[wsnet] cycles per step: from zero {r0['cycles_per_step']:.1f}, learned {r2['cycles_per_step']:.1f}; "
          f"velocities vs zero start {max(rel[:3]):.2e}, pressure {rel[3]:.2e}; div_max {r2['div_max']:.2e}",
          flush=True)
    # both starts iterate to max|residual| <= 1e-12: the velocities agree to
    # the solve tolerance, the pressure (which sums phi / H_BAR over the
    # substeps) to 1e-6, as for the other warm starts
    assert max(rel[:3]) < 1e-8
    assert rel[3] < 1e-6
    assert r2["div_max"] < 1e-10
    inf.close()
