"""Skin-model MLP restated in numpy float64 -- TEST INFRASTRUCTURE ONLY (the checker).

The skin model has no counterpart in the reference; this is the mathematical definition the HIP
kernels (softbox_amd/csrc/sbx_mlp.hip) are checked against: 12 -> 64 -> 64 -> 2, exact GELU,
inputs standardised by stored mean/std and clamped to +-6, loss = mean over samples of sum_o 0.5 * m_o * (y - t)^2.
Parameter layout: W1[64][12] | b1[64] | W2[64][64] | b2[64] | W3[2][64] | b3[2] | xMean[12] | xStd[12].
Its backward pass is pinned by central finite differences (tests/test_oracle_skin.py).
"""
from __future__ import annotations

import math

import numpy as np

IN, H, OUT = 12, 64, 2
W1 = 0
B1 = W1 + H * IN
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + OUT * H
NTRAIN = B3 + OUT
XMEAN = NTRAIN
XSTD = XMEAN + IN
NPARAM = XSTD + IN
XCLAMP = 6.0   # standardised inputs are clamped to [-XCLAMP, XCLAMP]

_erf = np.vectorize(math.erf)


def unpack(w):
    w = np.asarray(w, np.float64)
    return (w[W1:B1].reshape(H, IN), w[B1:W2], w[W2:B2].reshape(H, H), w[B2:W3], w[W3:B3].reshape(OUT, H),
            w[B3:NTRAIN], w[XMEAN:XSTD], w[XSTD:NPARAM])


def gelu(a):
    return 0.5 * a * (1.0 + _erf(a / math.sqrt(2.0)))


def gelu_grad(a):
    return 0.5 * (1.0 + _erf(a / math.sqrt(2.0))) + a * np.exp(-0.5 * a * a) / math.sqrt(2.0 * math.pi)


def forward(w, x):
    w1, b1, w2, b2, w3, b3, mu, sd = unpack(w)
    xs = np.clip((np.asarray(x, np.float64) - mu) / sd, -XCLAMP, XCLAMP)
    a1 = xs @ w1.T + b1
    h1 = gelu(a1)
    a2 = h1 @ w2.T + b2
    h2 = gelu(a2)
    return h2 @ w3.T + b3, (xs, a1, h1, a2, h2)


def loss_grad(w, x, t, m=(1.0, 1.0)):
    """Loss and gradient with respect to every parameter (the standardisation entries get 0)."""
    w1, b1, w2, b2, w3, b3, mu, sd = unpack(w)
    y, (xs, a1, h1, a2, h2) = forward(w, x)
    B = y.shape[0]
    m = np.asarray(m, np.float64)
    e = y - np.asarray(t, np.float64)
    loss = float((0.5 * m * e * e).sum() / B)
    dy = m * e / B
    g = np.zeros(NPARAM)
    g[W3:B3] = (dy.T @ h2).ravel()
    g[B3:NTRAIN] = dy.sum(0)
    da2 = (dy @ w3) * gelu_grad(a2)
    g[W2:B2] = (da2.T @ h1).ravel()
    g[B2:W3] = da2.sum(0)
    da1 = (da2 @ w2) * gelu_grad(a1)
    g[W1:B1] = (da1.T @ xs).ravel()
    g[B1:W2] = da1.sum(0)
    return loss, g


def init(rng: np.random.Generator, mean=None, std=None) -> np.ndarray:
    """He-style initial weights (float32), zero biases, given input standardisation."""
    w = np.zeros(NPARAM, np.float32)
    w[W1:B1] = rng.standard_normal(H * IN) * math.sqrt(2.0 / IN)
    w[W2:B2] = rng.standard_normal(H * H) * math.sqrt(2.0 / H)
    w[W3:B3] = rng.standard_normal(OUT * H) * math.sqrt(1.0 / H)
    w[XMEAN:XSTD] = 0.0 if mean is None else mean
    w[XSTD:NPARAM] = 1.0 if std is None else std
    return w
