"""Pins the FP64 CNN oracle (CPU only): conv vs scipy, backward vs finite differences."""
import numpy as np
import pytest
import scipy.ndimage

from oracle import cnn


def test_conv3_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(5, 6, 7, 3))
    w = rng.normal(size=(27, 4, 3))
    y = cnn.conv3(x, w)
    for co in range(4):
        exp = np.zeros(x.shape[:3])
        for ci in range(3):
            k = np.zeros((3, 3, 3))
            for t, (dx, dy, dz) in enumerate(cnn.TAPS):
                k[dx + 1, dy + 1, dz + 1] = w[t, co, ci]
            exp += scipy.ndimage.correlate(x[..., ci], k, mode="constant", cval=0.0)
        np.testing.assert_allclose(y[..., co], exp, atol=1e-12)


def test_conv3_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3, 4, 5, 2))
    w = rng.normal(size=(27, 3, 2))
    g = rng.normal(size=(3, 4, 5, 3))          # L = sum(g * conv3(x, w))
    dx = cnn.conv3_input_grad(g, w)
    dw = cnn.conv3_weight_grad(x, g)
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (2, 3, 4, 1), (1, 2, 2, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += eps
        xm[idx] -= eps
        fd = ((g * cnn.conv3(xp, w)).sum() - (g * cnn.conv3(xm, w)).sum()) / (2 * eps)
        assert dx[idx] == pytest.approx(fd, rel=1e-6, abs=1e-8)
    for idx in [(0, 0, 0), (13, 2, 1), (26, 1, 0)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += eps
        wm[idx] -= eps
        fd = ((g * cnn.conv3(x, wp)).sum() - (g * cnn.conv3(x, wm)).sum()) / (2 * eps)
        assert dw[idx] == pytest.approx(fd, rel=1e-6, abs=1e-8)


def test_network_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    params = cnn.init_params(rng, cin0=4, in_used=3, scale=1.0)
    x = rng.normal(size=(3, 3, 4, 4))
    dp = rng.uniform(0.5, 1.5, size=(3, 3, 4))
    dt = dp * rng.uniform(0.8, 1.2, size=dp.shape)
    m = (rng.random(dp.shape) > 0.2).astype(float)
    r, cache = cnn.forward(params, x)
    loss, dr = cnn.loss_and_grad_r(r, dp, dt, m)
    grads = cnn.backward(params, cache, dr)
    eps = 1e-6

    def L(ps):
        rr, _ = cnn.forward(ps, x)
        return cnn.loss_and_grad_r(rr, dp, dt, m)[0]

    for layer, which, idx in [(0, 0, (13, 5, 1)), (2, 0, (4, 7, 9)), (4, 1, (3,)), (5, 0, (13, 0, 5)),
                              (5, 1, (0,)), (1, 0, (26, 31, 31)), (3, 1, (17,))]:
        ps = [(w.copy(), b.copy()) for w, b in params]
        pm = [(w.copy(), b.copy()) for w, b in params]
        ps[layer][which][idx] += eps
        pm[layer][which][idx] -= eps
        fd = (L(ps) - L(pm)) / (2 * eps)
        assert grads[layer][which][idx] == pytest.approx(fd, rel=1e-5, abs=1e-9), (layer, which, idx)


def test_bf16_round_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.140625], dtype=np.float32)
    got = cnn.bf16_round(vals)
    np.testing.assert_array_equal(got, np.array([1.0, 1.0, 1.0 + 2 ** -6, -2.5, 3.140625], dtype=np.float32))


def test_forward_stored_with_identity_store_is_forward():
    rng = np.random.default_rng(12)
    params = cnn.init_params(rng)
    params = [(w, rng.normal(scale=0.05, size=32)) for w, _ in params]
    x = rng.normal(size=(5, 4, 6, 16))
    x[..., 3:] = 0.0
    r, _ = cnn.forward(params, x)
    np.testing.assert_allclose(cnn.forward_stored(params, x, store=lambda a: a), r, rtol=1e-12, atol=1e-12)
    # BF16 storage changes the result by about BF16 rounding, not more
    rb = cnn.forward_stored(params, x)
    assert 0 < np.abs(rb - r).max() <= 3e-2 * np.abs(r).max()
