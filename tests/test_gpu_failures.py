"""Failure modes of the reference trainer, forced through the inputs (every kernel runs to
its normal end): the cond(D) > 1e18 stop (tensorsvm-train.cu:1653-1658), a failed
capacitance Cholesky returning the current iterate (1709-1714), no free support vectors
giving a NaN bias (372), gamma = 0 giving a rank-1 kernel, a rank-deficient sketch, k > n,
and MAX_MPC_ITER = 1 (62).  Each is checked against the oracle where it has a counterpart."""
import numpy as np
import pytest

import oracle_lib as O
from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu

TSVM_CONVERGED, TSVM_ILL_COND, TSVM_MAXIT, TSVM_CHOL_FAIL = 0, 1, 2, 3


def _separable(n=400, k=8, sep=2.0, seed=5):
    rng = np.random.default_rng(seed)
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    U = rng.normal(size=(n, k)) * 0.3 + sep * y[:, None] * np.ones((1, k)) / np.sqrt(k)
    return (y[:, None] * U).astype(np.float32), y.astype(np.float32)


def test_ipm_ill_conditioned_stop(ctx):
    """Separable data and a gap tolerance below reach: D's spread passes 1e18 and both
    solvers stop with the ill-conditioning reason at the same iteration."""
    import tensorsvm_amd as T
    Z, y = _separable()
    C = 1e3
    al, xi, s, info = ctx.ipm(Z, y, C, stop=T.capi.STOP_GAP, tol=1e-14, maxit=200)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y.astype(np.float64), C, stop="gap", tol=1e-14,
                            maxit=200)
    assert io.reason == TSVM_ILL_COND
    assert info.reason == TSVM_ILL_COND, info.as_dict()
    assert info.iters == io.iters
    assert info.condD > 1e18
    assert np.abs(al - Xo).max() <= 1e-6 * C


def test_ipm_cholesky_failure_returns_iterate(ctx):
    """A NaN in the factor makes the capacitance's first pivot NaN: the solve stops with the
    Cholesky-failure reason before its first update (the reference's silent return)."""
    import tensorsvm_amd as T
    Z, y = _separable(sep=0.5)
    Z[17, 3] = np.nan
    al, xi, s, info = ctx.ipm(Z, y, 4.0, precision=T.PREC_FP32)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y.astype(np.float64), 4.0)
    assert io.reason == TSVM_CHOL_FAIL and io.iters == 0
    assert info.reason == TSVM_CHOL_FAIL and info.iters == 0, info.as_dict()


def test_ipm_one_iteration(ctx):
    """maxit = 1 (MAX_MPC_ITER = 1): one predictor-corrector step, the iteration-limit
    reason, and the oracle's iterate after the same step."""
    Z, y = _separable(sep=0.5)
    C = 4.0
    al, xi, s, info = ctx.ipm(Z, y, C, maxit=1)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y.astype(np.float64), C, maxit=1)
    assert info.reason == io.reason == TSVM_MAXIT
    assert info.iters == io.iters == 1
    assert np.abs(al - Xo).max() <= 1e-6 * C
    assert np.abs(xi - Xio).max() <= 1e-6 * C


def test_train_no_free_sv_nan_bias(ctx):
    """C = 1e-6: every multiplier sits at or below 1e-6, so there are no support vectors in
    the reference's sense and its bias divides by zero (372): NaN, as the oracle's."""
    import tensorsvm_amd as T
    X, y = gaussian_mixture(1000, 6, 3)
    k, C = 16, 1e-6
    om = O.gaussrand(1000 * k)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=0.1, precision=T.PREC_FP32, omega=om)
    U, _, _ = O.lra(X, y, k, 0.1, om.reshape(-1, k))
    alo, _, _, io = O.mpc(U, y.astype(np.float64), C)
    bo, cnt, _ = O.bias_rbf(alo, y, U, C)
    assert cnt == 0 and np.isnan(bo)
    assert info.nsv == 0 and info.nfree == 0 and info.b_samples == 0
    assert np.isnan(info.b)


def test_gamma_zero_rank_one_kernel(ctx):
    """gamma = 0: K = 1 1^T, so K_Y = y y^T has rank 1; the factor reproduces it (one
    eigenvalue n, the rest clipped or at rounding level).  With 600 / 400 labels the QP's
    optimum is -2 * 400 * C (every minority multiplier at C); the solver reaches the
    oracle's value.  (With balanced labels the reference's initial point 1^T Q 1 / |Q 1|^2,
    1535-1552, is 0 / 0 for an exactly rank-1 Q.)"""
    import tensorsvm_amd as T
    X, y = gaussian_mixture(1200, 8, 4)
    keep = np.sort(np.concatenate([np.flatnonzero(y > 0)[:600], np.flatnonzero(y < 0)[:400]]))
    X, y = X[keep], y[keep]
    n, k = 1000, 16
    om = O.gaussrand(n * k)
    U, info = ctx.lra(X, y, k, 0.0, T.PREC_FP32, omega=om)
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    yy = np.outer(y, y).astype(np.float64)
    assert np.abs(G - yy).max() <= 1e-5 * n
    assert info.realk >= 1
    assert abs(info.eig_max - n) <= 1e-4 * n
    Uo, wo, _ = O.lra(X, y, k, 0.0, om.reshape(n, k))
    assert np.abs(Uo @ Uo.T - yy).max() <= 1e-9 * n
    al, _, ti = ctx.train(X, y, k=k, C=1.0, gamma=0.0, precision=T.PREC_FP32, omega=om)
    _, _, _, io = O.mpc(y[:, None].astype(np.float64), y.astype(np.float64), 1.0)
    assert io.reason == 0 and abs(io.pobj + 800.0) <= 1e-6 * 800
    assert ti.ipm.reason == 0, ti.ipm.as_dict()
    assert abs(ti.ipm.pobj - io.pobj) <= 1e-6 * abs(io.pobj)
    assert np.all((al >= -1e-9) & (al <= 1.0 + 1e-9))


def test_rank_deficient_sketch(ctx):
    """Omega with zero columns: range(K_Y Omega) has dimension r < k, the basis Gram is
    singular and the factor keeps the r-dimensional projection P K_Y P (a pseudo-inverse of
    the Gram instead of its Cholesky), clipping the k - r empty directions."""
    import tensorsvm_amd as T
    from scipy.spatial.distance import cdist
    X, y = gaussian_mixture(800, 6, 6)
    n, k, gamma = 800, 24, 0.1
    om = O.gaussrand(n * k).reshape(n, k)
    om[:, 5] = 0.0
    om[:, 11] = 0.0
    U, info = ctx.lra(X, y, k, gamma, T.PREC_FP32, omega=om)
    yd = y.astype(np.float64)
    Ky = yd[:, None] * np.exp(-gamma * cdist(X.astype(np.float64), X.astype(np.float64),
                                              "sqeuclidean")) * yd[None, :]
    keep = [c for c in range(k) if c not in (5, 11)]
    Q, _ = np.linalg.qr(Ky @ om[:, keep].astype(np.float64))
    w, V = np.linalg.eigh(Q.T @ Ky @ Q)
    F = Q @ (V * np.sqrt(np.maximum(w, 0)))
    Go = F @ F.T
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    assert np.linalg.norm(G - Go) / np.linalg.norm(Go) < 1e-4
    assert info.realk <= k - 2


def test_rank_above_rows_rejected(ctx):
    """k > n (and k above the solver's limit) is an argument error before any work."""
    import tensorsvm_amd as T
    X, y = gaussian_mixture(100, 4, 1)
    with pytest.raises(T.capi.TsvmError, match="status 1"):
        ctx.train(X, y, k=101, C=1.0, gamma=0.1)
    X2, y2 = gaussian_mixture(2000, 4, 1)
    with pytest.raises(T.capi.TsvmError, match="status 1"):
        ctx.train(X2, y2, k=1100, C=1.0, gamma=0.1)
