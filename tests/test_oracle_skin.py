"""Pins of the skin-model oracle (CPU only): the numpy backward pass against central finite
differences of its own loss, GELU against its closed-form limits, and the loss at a known point."""
import math

import numpy as np

from oracle import skin_mlp as M


def test_parameter_layout():
    assert M.NTRAIN == 64 * 12 + 64 + 64 * 64 + 64 + 2 * 64 + 2
    assert M.NPARAM == M.NTRAIN + 24


def test_gelu_limits_and_slope():
    a = np.array([-8.0, 0.0, 8.0, 1.0])
    g = M.gelu(a)
    assert abs(g[0]) < 1e-12 and g[1] == 0.0 and abs(g[2] - 8.0) < 1e-12
    assert abs(g[3] - 0.8413447460685429) < 1e-12          # Phi(1)
    h = 1e-6
    for x in (-2.0, -0.3, 0.0, 0.7, 3.0):
        fd = (M.gelu(np.array([x + h]))[0] - M.gelu(np.array([x - h]))[0]) / (2 * h)
        assert abs(fd - M.gelu_grad(np.array([x]))[0]) < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = M.init(rng, mean=rng.standard_normal(12), std=0.5 + rng.random(12)).astype(np.float64)
    w[M.B1:M.W2] = rng.standard_normal(64) * 0.1
    w[M.B2:M.W3] = rng.standard_normal(64) * 0.1
    x = rng.standard_normal((7, 12))
    t = rng.standard_normal((7, 2))
    m = (1.0, 0.3)
    _, g = M.loss_grad(w, x, t, m)
    idx = np.concatenate([rng.choice(np.arange(M.W1, M.B1), 12, replace=False),
                          np.arange(M.B1, M.B1 + 4), rng.choice(np.arange(M.W2, M.B2), 12, replace=False),
                          np.arange(M.B2, M.B2 + 4), np.arange(M.W3, M.NTRAIN)])
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (M.loss_grad(wp, x, t, m)[0] - M.loss_grad(wm, x, t, m)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    assert not g[M.XMEAN:].any()


def test_zero_weights_give_bias_output_and_closed_form_loss():
    w = np.zeros(M.NPARAM)
    w[M.XSTD:] = 1.0
    w[M.B3:M.NTRAIN] = (0.5, -2.0)
    x = np.ones((3, 12))
    y, _ = M.forward(w, x)
    assert np.allclose(y, [[0.5, -2.0]] * 3)
    t = np.array([[1.5, -2.0]] * 3)
    loss, g = M.loss_grad(w, x, t, (2.0, 1.0))
    assert math.isclose(loss, 0.5 * 2.0 * 1.0)              # per sample 0.5 * m0 * (0.5 - 1.5)^2
    assert math.isclose(g[M.B3], 2.0 * -1.0) and g[M.B3 + 1] == 0.0
