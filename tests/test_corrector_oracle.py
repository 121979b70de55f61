"""Pins the FP64 corrector oracle (oracle/corrector.py) by central finite
differences and by a direct per-pixel evaluation of one convolution."""
import numpy as np
import pytest

from oracle import corrector as cn


def test_param_count():
    assert cn.n_params() == 3 * 32 * 9 + 32 + 2 * (32 * 32 * 9 + 32) + 32 * 2 * 9 + 2 == 19970


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    H, W = 5, 7
    x = rng.standard_normal((H, W, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    y = cn.conv(x, w, b)
    for (h, v) in [(0, 0), (4, 6), (2, 3)]:
        want = b.copy()
        for dy in range(3):
            for dx in range(3):
                want += w[:, :, dy, dx] @ x[(h + dy - 1) % H, (v + dx - 1) % W]
        assert np.allclose(y[h, v], want, rtol=1e-13, atol=1e-13)


def test_gradient_finite_differences():
    rng = np.random.default_rng(1)
    H = W = 6
    theta = cn.init_params(3, last_scale=0.3)
    x = rng.standard_normal((H, W, 3))
    t = rng.standard_normal((H, W, 2))
    loss, g = cn.loss_and_grad(theta, x, t)
    idx = rng.choice(theta.size, 40, replace=False)
    # first and last weight and first bias of layer 0, the V channel's first
    # tap, the last layer's weights and biases
    idx = np.concatenate([idx, [0, 863, 864, 18, 19969, 19968, 19390]])
    for i in idx:
        e = np.zeros_like(theta)
        h = 1e-5 * max(1.0, abs(theta[i]))
        e[i] = h
        fd = (cn.loss_and_grad(theta + e, x, t)[0] - cn.loss_and_grad(theta - e, x, t)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_dgrad_is_adjoint_of_conv():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((5, 6, 3))
    dz = rng.standard_normal((5, 6, 4))
    w = rng.standard_normal((4, 3, 3, 3))
    lhs = (cn.conv(x, w, np.zeros(4)) * dz).sum()
    rhs = (x * cn.conv_dgrad(dz, w)).sum()
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_bf16_rounding():
    assert cn.round_bf16(1.0) == 1.0
    assert cn.round_bf16(1.0 + 2 ** -9) == 1.0        # tie to even
    assert cn.round_bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7
    assert cn.round_bf16(-3.14159) == pytest.approx(-3.140625)
