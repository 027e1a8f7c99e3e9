"""FP64 numpy restatement of the near-field learning problem (test infrastructure).

* label(x): the 16 x 16 sub-fragment quadrature of the pair kernel over the
  emitter's cell, minus the base rule (the unregularised kernel averaged over
  the cell's 2 x 2 Gauss-Legendre points, numpy's leggauss nodes), times
  e^{-ikr} / M and the target gain, rebuilt from the features alone (lengths in wavelengths,
  emitter frame).
* forward / gradient: the 10-64-64-2 tanh network and the gradient of
  L = sum_i |out_i - y_i|^2 / n, by explicit back-propagation.
"""
import numpy as np

NIN = 10  # network inputs
OFF_B1 = 64 * NIN
OFF_W2 = OFF_B1 + 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 2 * 64
NPAR = OFF_B3 + 2
TARGET_GAIN = 4096.0  # the library's kNfTargetGain


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[:OFF_B1].reshape(64, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64), p[OFF_B2:OFF_W3],
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


def gauss2(rv, nr, d, k):
    """The base rule: the kernel (no regularisation, cosine culls) at the 2 x 2
    Gauss-Legendre points of the d x d emitter cell, equally weighted."""
    x, _ = np.polynomial.legendre.leggauss(2)
    o = x * d / 2
    acc = 0.0
    for oy in o:
        for ox in o:
            v = np.array([rv[0] - ox, rv[1] - oy, rv[2]])
            rs = np.linalg.norm(v)
            css, cis = v[2] / rs, -(v @ nr) / rs
            if css >= 1e-6 and cis >= 1e-6:
                acc = acc + _g(rs, css, cis, k, 0.0)
    return acc / 4


def label(x):
    """Target of one feature vector (k = 2 pi, lambda = 1)."""
    k = 2 * np.pi
    r = x[0] + 1.0
    d = x[1] / 14.0 + 0.19
    u = np.array([x[2], x[3], (x[4] + 1.0) / 2.0], np.float64)
    nr = np.array(x[5:8], np.float64)
    rv = u * r
    h = d / 16
    o = (np.arange(16) + 0.5) * h - 0.5 * d
    ox, oy = np.meshgrid(o, o, indexing="xy")
    v = np.stack([rv[0] - ox, rv[1] - oy, np.full_like(ox, rv[2])], -1)
    rs = np.linalg.norm(v, axis=-1)
    css, cis = v[..., 2] / rs, -(v @ nr) / rs
    ok = (css >= 1e-6) & (cis >= 1e-6)
    Q = np.where(ok, _g(rs, css, cis, k, h), 0).sum() / 256
    K = gauss2(rv, nr, d, k)
    M = 0.9 * (k / (4 * np.pi * r) + 1 / r ** 2)
    t = TARGET_GAIN * (Q - K) * np.exp(-1j * k * r) / M
    return np.array([t.real, t.imag])
