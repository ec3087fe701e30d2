"""The oracle's restatement of the warm-start network (oracle/wsnet.py), pinned
on the CPU without the GPU: its gradient against central finite differences,
its upsampling against the adjoint identity, its strided convolution against
a brute-force sum, and the input preparation against its definition."""
import numpy as np
import pytest

from oracle import wsnet as W


def _rand_case(seed, n=2, h1=17, w1=17):
    rng = np.random.default_rng(seed)
    theta = W.he_init(seed)
    x = rng.standard_normal((n, h1, w1, W.CIN0))
    t0 = rng.standard_normal((n, 2 * h1 - 1, 2 * w1 - 1, 1))
    return theta, x, t0


def test_parameter_count_is_about_200k():
    n = W.nparams()
    print("warm-start network parameters:", n)
    assert 180_000 < n < 260_000


def test_conv3_matches_a_direct_sum():
    rng = np.random.default_rng(1)
    for s in (1, 2):
        x = rng.standard_normal((2, 7, 9, 4))
        w = rng.standard_normal((5, 4, 9))
        b = rng.standard_normal(5)
        y = W.conv3(x, w, b, s, lambda a: a)
        ho, wo = (7 - 1) // s + 1, (9 - 1) // s + 1
        assert y.shape == (2, ho, wo, 5)
        ref = np.zeros_like(y)
        for n in range(2):
            for oy in range(ho):
                for ox in range(wo):
                    for co in range(5):
                        acc = b[co]
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                iy, ix = s * oy + dy, s * ox + dx
                                if 0 <= iy < 7 and 0 <= ix < 9:
                                    acc += w[co, :, 3 * (dy + 1) + dx + 1] @ x[n, iy, ix, :]
                        ref[n, oy, ox, co] = acc
        assert np.abs(y - ref).max() < 1e-12


def test_upsampling_adjoint_identity():
    rng = np.random.default_rng(2)
    c = rng.standard_normal((2, 5, 9, 3))
    g = rng.standard_normal((2, 9, 17, 3))
    lhs = (W.up(c, (9, 17)) * g).sum()
    rhs = (c * W.up_adj(g)).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # bilinear: reproduces linear functions exactly
    yy, xx = np.meshgrid(np.arange(5.0), np.arange(9.0), indexing="ij")
    lin = (2 * yy - 0.5 * xx)[None, :, :, None]
    fy, fx = np.meshgrid(np.arange(9.0) / 2, np.arange(17.0) / 2, indexing="ij")
    assert np.abs(W.up(lin, (9, 17))[0, :, :, 0] - (2 * fy - 0.5 * fx)).max() < 1e-14


def test_gradient_matches_finite_differences():
    theta, x, t0 = _rand_case(3)
    loss, g = W.loss_and_grad(theta, x, t0)
    rng = np.random.default_rng(4)
    # every layer's weights and biases, a few entries each
    idx, off = [], 0
    for _i, co, ci, t in W.param_layout():
        nw = co * ci * t
        idx += list(off + rng.choice(nw, 3, replace=False)) + [off + nw + int(rng.integers(co))]
        off += nw + co
    h = 1e-6
    worst = 0.0
    for k in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (W.loss_and_grad(tp, x, t0)[0] - W.loss_and_grad(tm, x, t0)[0]) / (2 * h)
        err = abs(fd - g[k]) / max(abs(fd), 1e-3 * np.abs(g).max())
        worst = max(worst, err)
    print(f"oracle network gradient vs central differences: loss {loss:.4f}, worst relative {worst:.2e} "
          f"over {len(idx)} parameters")
    assert worst < 1e-5


def test_bf16_rounding_is_nearest_even():
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1)
    assert W.bf16(np.array([1.0 + 2 ** -8]))[0] == 1.0
    assert W.bf16(np.array([1.0 + 3 * 2 ** -8]))[0] == 1.0 + 2 ** -6
    assert W.bf16(np.array([1.0 + 2 ** -8 + 2 ** -12]))[0] == 1.0 + 2 ** -7
    v = np.random.default_rng(5).standard_normal(1000)
    r = W.bf16(v)
    assert np.all(np.abs(r - v) <= 2 ** -8 * np.abs(v))


def test_prepare_inputs():
    rng = np.random.default_rng(6)
    rhs = rng.standard_normal((3, 33, 65))
    rhs[2] = 0.0
    prev = rng.standard_normal((3, 33, 65))
    kx = np.array([0.5, 1.0, 2.0])
    x, s = W.prepare(rhs, prev, kx, 2.0, 10.0)
    assert x.shape == (3, 17, 33, 4)
    assert s[0] == np.abs(rhs[0]).max() and s[2] == 0.0
    # full weighting of an interior point, then the scale
    r = rhs[0]
    want = (r[1:4, 3:6] * np.outer([0.25, 0.5, 0.25], [0.25, 0.5, 0.25])).sum() / s[0]
    assert abs(x[0, 1, 2, 0] - want) < 1e-14
    assert abs(x[1, 0, 0, 2] - 0.5) < 1e-15 and x[1, -1, 0, 3] == 1.0
    assert np.all(x[2] == 0.0)
    # the previous solution enters scaled by the H_BAR ratio
    x2, _ = W.prepare(rhs, prev, kx, 2.0, 10.0, ratio=0.25)
    np.testing.assert_allclose(x2[..., 1], 0.25 * x[..., 1], rtol=1e-15)
    assert np.array_equal(x2[..., 0], x[..., 0])


def test_target_and_guess_are_the_scaled_previous_solution_plus_the_output():
    """The target is what the network adds to ratio * prev (normalised), and
    a network whose output is zero guesses ratio * prev exactly."""
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((2, 33, 33))
    prev = rng.standard_normal((2, 33, 33))
    phi = 0.4 * prev + 0.01 * rng.standard_normal((2, 33, 33))
    s = np.abs(rhs).max(axis=(1, 2))
    t = W.target(phi, prev, 0.4, s, 3.0)
    np.testing.assert_allclose(t * (s * 3.0)[:, None, None] + 0.4 * prev, phi, rtol=1e-12)
    theta = np.zeros(W.nparams())
    g = W.guess(theta, rhs, prev, np.array([0.5, 1.0]), 2.0, 3.0, ratio=0.4)
    np.testing.assert_array_equal(g, 0.4 * prev)
