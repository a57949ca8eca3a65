"""Pins of the U-Net oracle (oracle/unet_ref.py) by what the mathematics fixes."""
import os
import sys

import numpy as np
import pytest
from scipy.signal import correlate2d

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import unet_ref as U  # noqa: E402


def test_conv3x3_matches_scipy_correlate():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 7, 5, 3))
    w = rng.standard_normal((9, 4, 3))
    y = U.conv3x3(x, w)
    for n in range(2):
        for co in range(4):
            ref = np.zeros((7, 5))
            for ci in range(3):
                k = w[:, co, ci].reshape(3, 3)  # k[dx+1, dz+1]
                ref += correlate2d(x[n, :, :, ci], k, mode="same", boundary="fill")
            np.testing.assert_allclose(y[n, :, :, co], ref, rtol=1e-12, atol=1e-12)


def test_conv_dgrad_wgrad_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 4, 3))
    w = rng.standard_normal((9, 5, 3))
    dy = rng.standard_normal((2, 6, 4, 5))
    lhs = (U.conv3x3(x, w) * dy).sum()
    assert abs(lhs - (x * U.conv3x3_dgrad(dy, w)).sum()) < 1e-10 * abs(lhs)
    assert abs(lhs - (w * U.conv3x3_wgrad(x, dy)).sum()) < 1e-10 * abs(lhs)


def test_pool_upsample_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 8, 6, 3))
    y = rng.standard_normal((2, 4, 3, 3))
    assert abs((U.pool2(x) * y).sum() - (x * U.pool2_bwd(y)).sum()) < 1e-12
    assert abs((U.up2(y) * x).sum() - (y * U.up2_bwd(x)).sum()) < 1e-12
    # average of a constant is the constant; nearest upsampling replicates
    assert np.allclose(U.pool2(np.full((1, 4, 4, 1), 3.0)), 3.0)
    assert U.up2(y).shape == (2, 8, 6, 3) and np.allclose(U.up2(y)[:, ::2, 1::2], y)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3])
    r = U.bf16(v)
    # 8 mantissa bits: 1+2^-8 is a tie -> even (1.0); 1+3*2^-8 tie -> 1+2^-6 (even)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0 and r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


@pytest.mark.parametrize("levels,base", [(1, 3), (2, 2), (3, 2)])
def test_unet_gradient_finite_difference(levels, base):
    rng = np.random.default_rng(4)
    n, h, w = 2, 8, 8
    theta = U.init_params(levels, base, seed=5)
    theta += 0.05 * rng.standard_normal(theta.size)  # nonzero biases: ReLU kinks away from 0
    x = rng.standard_normal((n, h, w))
    t = rng.standard_normal((n, h, w))
    loss, g, _ = U.loss_grad(theta, x, t, levels, base)
    idx = rng.choice(theta.size, size=min(60, theta.size), replace=False)
    # cover every parameter block at least once (first element of each)
    o = 0
    for _, ci, co, k in U.layers(levels, base):
        idx = np.append(idx, [o, o + (9 if k == 3 else 1) * co * ci])
        o += (9 if k == 3 else 1) * co * ci + co
    eps = 1e-6
    for i in np.unique(idx):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (U.loss_grad(tp, x, t, levels, base)[0] - U.loss_grad(tm, x, t, levels, base)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_unet_shapes_and_param_count():
    assert U.num_params(4, 32) == U.init_params(4, 32).size
    theta = U.init_params(4, 8, seed=1)
    y, _ = U.forward(theta, np.zeros((1, 16, 8)), 4, 8)
    assert y.shape == (1, 16, 8)
