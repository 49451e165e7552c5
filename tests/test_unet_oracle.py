"""Pins the FP64 U-Net oracle (oracle/unet.cpp) on the CPU:
closed-form outputs of hand-built weight sets (tap offsets, channel routing,
pooling/up-sampling), the masked MSE, and the backward pass against central
finite differences of the forward, per layer."""
import numpy as np
import pytest

import unet_params as U


def test_param_count(oracle):
    assert oracle.unet_param_count() == U.COUNT == 321745


@pytest.mark.parametrize("d", [(1, 0, 0), (0, -1, 0), (0, 0, 1), (-1, 1, -1), (0, 0, 0)])
def test_shift_taps(oracle, d):
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2, 8, 4, 12, 4))
    for ch in (0, 3):
        out, _ = oracle.unet_forward(U.shift_params(*d, channel=ch), x)
        want = np.maximum(U.shifted(x[..., ch], *d), 0.0)
        np.testing.assert_allclose(out, want, rtol=0, atol=1e-14)


def test_coarse_path_block_means(oracle):
    rng = np.random.default_rng(4)
    x = rng.uniform(0.0, 1.0, size=(1, 8, 12, 16, 4))
    out, _ = oracle.unet_forward(U.coarse_path_params(), x)
    np.testing.assert_allclose(out, U.block_mean_up(x[..., 0], 4), rtol=0, atol=1e-14)


def test_bias_only(oracle):
    # every weight zero: d0 = relu(b9) at every voxel, z = head_b + head_w . relu(b9)
    p = np.zeros(U.COUNT)
    U.conv_bias(p, 9)[:] = np.linspace(-1.0, 1.0, 16)
    hw, hb = U.head(p)
    hw[:] = np.arange(16) * 0.1
    hb[:] = 0.5
    x = np.random.default_rng(5).normal(size=(1, 4, 4, 4, 4))
    out, _ = oracle.unet_forward(p, x)
    want = 0.5 + np.dot(hw, np.maximum(U.conv_bias(p, 9), 0.0))
    np.testing.assert_allclose(out, want, rtol=1e-15, atol=1e-15)


def test_masked_mse(oracle):
    rng = np.random.default_rng(6)
    p = U.random_params(7)
    x = rng.normal(size=(1, 8, 8, 8, 4))
    t = rng.normal(size=(1, 8, 8, 8))
    m = (rng.uniform(size=t.shape) < 0.6).astype(float)
    out, loss = oracle.unet_forward(p, x, t, m)
    assert loss == pytest.approx(np.sum(m * (out - t) ** 2) / np.sum(m), rel=1e-12)
    _, loss_all = oracle.unet_forward(p, x, t)
    assert loss_all == pytest.approx(np.mean((out - t) ** 2), rel=1e-12)


def _layer_dirs(rng):
    """Random unit directions supported on one block of parameters each."""
    dirs = []
    for k in range(10):
        v = np.zeros(U.COUNT)
        w = U.conv_weights(v, k)
        w[:] = rng.normal(size=w.shape)
        if k == 0:
            w[:, :, 4:] = 0.0
        dirs.append((f"w{k}", v))
        v = np.zeros(U.COUNT)
        U.conv_bias(v, k)[:] = rng.normal(size=U.COUT[k])
        dirs.append((f"b{k}", v))
    v = np.zeros(U.COUNT)
    v[U.OFF_HEAD:] = rng.normal(size=17)
    dirs.append(("head", v))
    return dirs


def test_gradients_match_finite_differences(oracle):
    rng = np.random.default_rng(8)
    p = U.random_params(9)
    x = rng.normal(size=(2, 8, 8, 4, 4))
    t = rng.normal(size=(2, 8, 8, 4))
    m = (rng.uniform(size=t.shape) < 0.7).astype(float)
    loss, g = oracle.unet_gradients(p, x, t, m)
    _, loss_f = oracle.unet_forward(p, x, t, m)
    assert loss == pytest.approx(loss_f, rel=1e-14)
    # input channels 4..7 of c0 never see data: their gradient is exactly 0
    assert np.all(U.conv_weights(g, 0)[:, :, 4:] == 0.0)
    h = 1e-6
    for name, v in _layer_dirs(rng):
        v /= np.linalg.norm(v)
        _, lp = oracle.unet_forward(p + h * v, x, t, m)
        _, lm = oracle.unet_forward(p - h * v, x, t, m)
        fd = (lp - lm) / (2 * h)
        an = float(np.dot(g, v))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (name, fd, an)


def test_gradient_of_single_weights(oracle):
    """Element-wise checks on individual taps, catching a transposed
    [cout][cin] or flipped tap order that a random direction could average out."""
    rng = np.random.default_rng(10)
    p = U.random_params(11)
    x = rng.normal(size=(1, 8, 8, 8, 4))
    t = rng.normal(size=(1, 8, 8, 8))
    _, g = oracle.unet_gradients(p, x, t)
    h = 1e-6
    picks = [(0, 0, 3, 2), (0, 26, 15, 0), (1, 4, 2, 9), (5, 22, 63, 1), (6, 7, 31, 60), (9, 18, 5, 12)]
    for k, tap, co, ci in picks:
        i = U.OFF_W[k] + (tap * U.COUT[k] + co) * U.CIN[k] + ci
        e = np.zeros(U.COUNT)
        e[i] = h
        _, lp = oracle.unet_forward(p + e, x, t)
        _, lm = oracle.unet_forward(p - e, x, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 2e-5 * max(1e-3, abs(g[i])) + 1e-9, (k, tap, co, ci, fd, g[i])


def test_mutation_is_caught(oracle):
    """A planted mistake in the expected tap ordering (x and z swapped) fails
    the shift check: the check can tell the axes apart."""
    x = np.random.default_rng(12).normal(size=(1, 8, 8, 8, 4))
    out, _ = oracle.unet_forward(U.shift_params(1, 0, 0), x)
    wrong = np.maximum(U.shifted(x[..., 0], 0, 0, 1), 0.0)
    assert np.max(np.abs(out - wrong)) > 0.1
