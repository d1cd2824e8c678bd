"""Seeded synthetic inputs shaped like the benchmark configurations (tests/bench only)."""
import numpy as np


def gaussian_mixture(n, d, seed):
    """x = 0.5 y 1 + N(0, I), labels +-1 balanced (seeded shuffle)."""
    rng = np.random.default_rng(seed)
    y = np.ones(n, dtype=np.float32)
    y[n // 2:] = -1
    rng.shuffle(y)
    X = (0.5 * y[:, None] + rng.standard_normal((n, d))).astype(np.float32)
    return X, y


def covtype_like(n, seed, gamma_r=1.0, flips=0.05):
    """10 uniform[0,1] columns + 44 binary columns with exactly 2 ones per row; labels are
    the sign of a random 64-centre RBF expansion (median split) with seeded flips."""
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 54), dtype=np.float32)
    X[:, :10] = rng.random((n, 10), dtype=np.float32)
    ones = np.argsort(rng.random((n, 44)), axis=1)[:, :2]
    X[np.arange(n)[:, None], 10 + ones] = 1.0
    cidx = rng.choice(n, 64, replace=False)
    Cn = X[cidx].astype(np.float64)
    w = rng.standard_normal(64)
    f = np.zeros(n)
    for s in range(0, n, 65536):
        xs = X[s:s + 65536].astype(np.float64)
        d2 = (xs * xs).sum(1)[:, None] + (Cn * Cn).sum(1)[None, :] - 2 * xs @ Cn.T
        f[s:s + 65536] = np.exp(-gamma_r * np.maximum(d2, 0)) @ w
    y = np.where(f > np.median(f), 1.0, -1.0).astype(np.float32)
    flip = rng.random(n) < flips
    y[flip] = -y[flip]
    return X, y


def scaled_mixture(n, d, seed, flips=0.03):
    """Differentiation-form data: per-feature variances 10^U[-1, 1], x ~ N(0, diag s2);
    labels = sign(w.x + sum_{j<32} v_j exp(-|x - c_j|^2 / d)) with seeded flips.
    Returns X, y and the proposed gamma = 1 / sum s2."""
    rng = np.random.default_rng(seed)
    s2 = 10.0 ** rng.uniform(-1.0, 1.0, d)
    X = (rng.standard_normal((n, d)) * np.sqrt(s2)).astype(np.float32)
    w = rng.standard_normal(d) / np.sqrt(s2)
    cidx = rng.choice(n, 32, replace=False)
    Cn = X[cidx].astype(np.float64)
    v = rng.standard_normal(32) * 2.0
    f = X.astype(np.float64) @ w
    for s in range(0, n, 65536):
        xs = X[s:s + 65536].astype(np.float64)
        d2 = (xs * xs).sum(1)[:, None] + (Cn * Cn).sum(1)[None, :] - 2 * xs @ Cn.T
        f[s:s + 65536] += np.exp(-np.maximum(d2, 0) / d) @ v
    y = np.where(f > 0, 1.0, -1.0).astype(np.float32)
    flip = rng.random(n) < flips
    y[flip] = -y[flip]
    return X, y, float(1.0 / s2.sum())

