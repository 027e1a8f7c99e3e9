"""FP64 numpy restatement of the near-field learning problem (test infrastructure).

* label(x): the 16 x 16 sub-fragment quadrature of the pair kernel over the
  emitter's cell, minus the unregularised point value, times e^{-ikr} / M,
  rebuilt from the features alone (lengths in wavelengths, emitter frame).
* forward / gradient: the 9-64-64-2 tanh network and the gradient of
  L = sum_i |out_i - y_i|^2 / n, by explicit back-propagation.
"""
import numpy as np

OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, NPAR = 576, 640, 4736, 4800, 4928, 4930


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[:OFF_B1].reshape(64, 9), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3].reshape(2, 64), p[OFF_B3:NPAR])


def forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h1, h2, h2 @ W3.T + b3


def gradient(p, X, Y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, np.float64)
    n = len(X)
    h1, h2, out = forward(p, X)
    g3 = 2.0 * (out - np.asarray(Y, np.float64)) / n
    d2 = (g3 @ W3) * (1 - h2 ** 2)
    d1 = (d2 @ W2) * (1 - h1 ** 2)
    return np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), (g3.T @ h2).ravel(),
                           g3.sum(0)])


def _g(r, cs, ci, k, reg):
    G = np.where(r < reg, 1j * k / (4 * np.pi), -np.exp(1j * k * r) / (4 * np.pi) * (1j * k / r) + 1 / r ** 2)
    return cs * ci * 0.9 * G


def label(x):
    """Target of one feature vector (k = 2 pi, lambda = 1)."""
    k = 2 * np.pi
    r = x[0] + 1.0
    d = x[1] / 14.0 + 0.19
    u = np.array([x[2], x[3], (x[4] + 1.0) / 2.0], np.float64)
    nr = np.array(x[5:8], np.float64)
    rv = u * r
    cs, ci = u[2], -u @ nr
    h = d / 16
    o = (np.arange(16) + 0.5) * h - 0.5 * d
    ox, oy = np.meshgrid(o, o, indexing="xy")
    v = np.stack([rv[0] - ox, rv[1] - oy, np.full_like(ox, rv[2])], -1)
    rs = np.linalg.norm(v, axis=-1)
    css, cis = v[..., 2] / rs, -(v @ nr) / rs
    ok = (css >= 1e-6) & (cis >= 1e-6)
    Q = np.where(ok, _g(rs, css, cis, k, h), 0).sum() / 256
    K = _g(r, cs, ci, k, 0.0)
    M = 0.9 * (k / (4 * np.pi * r) + 1 / r ** 2)
    t = (Q - K) * np.exp(-1j * k * r) / M
    return np.array([t.real, t.imag])
