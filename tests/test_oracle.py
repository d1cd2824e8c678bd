"""CPU oracle pinned against golden vectors, library results and closed forms (no GPU)."""
import os

import numpy as np
import pytest
from scipy.spatial.distance import cdist

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
G = os.path.join(HERE, "golden")


def test_glibc_rand_and_gaussrand_golden():
    g = np.load(os.path.join(G, "glibc_rand.npz"))
    n = g["rand"].size
    assert np.array_equal(O.glibc_rand(n), g["rand"])
    assert np.array_equal(O.gaussrand(n).view(np.uint32), g["gauss_bits"])


def test_gaussrand_distribution():
    v = O.gaussrand(400_000).astype(np.float64)
    assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01


def test_kernel_product_closed_form():
    rng = np.random.default_rng(0)
    X1, X2 = rng.normal(size=(37, 6)).astype(np.float32), rng.normal(size=(53, 6)).astype(np.float32)
    y1 = np.sign(rng.normal(size=37)).astype(np.float32)
    y2 = np.sign(rng.normal(size=53)).astype(np.float32)
    A = rng.normal(size=(53, 5))
    K = np.exp(-0.3 * cdist(X1.astype(np.float64), X2.astype(np.float64), "sqeuclidean"))
    ref = (y1[:, None] * K * y2[None, :]) @ A
    assert np.allclose(O.kernel_product(X1, y1, X2, y2, A, 0.3), ref, rtol=1e-12, atol=1e-12)


def test_householder_q():
    rng = np.random.default_rng(1)
    W = rng.normal(size=(200, 12))
    Q = O.householder_q(W)
    assert np.allclose(Q.T @ Q, np.eye(12), atol=1e-12)
    Qn, _ = np.linalg.qr(W)
    assert np.allclose(Q @ Q.T, Qn @ Qn.T, atol=1e-12)  # same range


def test_sym_eig_vs_numpy():
    rng = np.random.default_rng(2)
    B = rng.normal(size=(20, 20))
    A = B + B.T
    w, V = O.sym_eig(A)
    wn = np.linalg.eigvalsh(A)
    assert np.allclose(w, wn, atol=1e-10)
    assert np.allclose(V @ np.diag(w) @ V.T, A, atol=1e-10)


def test_lra_full_rank_is_exact():
    """k = n (distinct points, K_Y positive definite): U U^T = K_Y."""
    rng = np.random.default_rng(3)
    n = 60
    X = rng.normal(size=(n, 4)).astype(np.float32)
    y = np.sign(rng.normal(size=n)).astype(np.float32)
    om = O.gaussrand(n * n).reshape(n, n)
    U, w, realk = O.lra(X, y, n, 0.4, om)
    K = y[:, None] * np.exp(-0.4 * cdist(X.astype(np.float64), X.astype(np.float64), "sqeuclidean")) * y[None, :]
    assert realk == n
    assert np.allclose(U @ U.T, K, atol=1e-9)


def test_lra_low_rank_is_projection():
    """U U^T = P K_Y P with P the projector on range(K_Y Omega) (k < n)."""
    rng = np.random.default_rng(4)
    n, k = 120, 10
    X = rng.normal(size=(n, 3)).astype(np.float32)
    y = np.sign(rng.normal(size=n)).astype(np.float32)
    om = O.gaussrand(n * k).reshape(n, k)
    U, w, realk = O.lra(X, y, k, 0.5, om)
    K = y[:, None] * np.exp(-0.5 * cdist(X.astype(np.float64), X.astype(np.float64), "sqeuclidean")) * y[None, :]
    Qn, _ = np.linalg.qr(K @ om.astype(np.float64))
    P = Qn @ Qn.T
    PKP = P @ K @ P
    ev, V = np.linalg.eigh((PKP + PKP.T) / 2)
    ref = (V * np.maximum(ev, 0)) @ V.T
    assert np.allclose(U @ U.T, ref, atol=1e-8)
    fk, fr = O.fnorm_res(X, y, 0.5, U)
    assert np.isclose(fk, np.linalg.norm(K)) and np.isclose(fr, np.linalg.norm(K - U @ U.T))


@pytest.mark.parametrize("name", ["a", "b"])
def test_mpc_vs_sklearn_svc(name):
    f = np.load(os.path.join(G, f"ipm_svc_{name}.npz"))
    C = float(f["C"])
    X, Xi, S, info = O.mpc(f["Z"], f["y"], C)
    assert info.reason == 0
    assert np.abs(X - f["alpha"]).max() <= 1e-5 * C
    assert abs(info.y + float(f["b"])) <= 1e-4 * (1 + abs(float(f["b"])))  # y = rho = -b
    # KKT at exit: a^T alpha = 0, box, complementarity = 2 n mu
    assert abs(np.dot(f["y"], X)) < 1e-6
    assert X.min() > 0 and X.max() < C


def test_smw_solve_vs_dense():
    rng = np.random.default_rng(5)
    n, k = 150, 7
    Z = rng.normal(size=(n, k))
    Dv = np.exp(rng.uniform(-3, 3, n))
    b = rng.normal(size=n)
    x = O.smw_solve(Z, Dv, b)
    assert np.allclose((np.diag(Dv) + Z @ Z.T) @ x, b, atol=1e-9)


def test_newton_step_solves_kkt_jacobian():
    """NewtonStep + SMWSolve (1794-1840) against a dense solve of the (3n+1) KKT Jacobian
    [-Q a I -I; a^T 0 0 0; S 0 X 0; -Xi 0 0 C-X] d = r (the system in the comment above
    mpc, tensorsvm-train.cu:1497-1501)."""
    rng = np.random.default_rng(6)
    n, k, C = 30, 4, 2.0
    a = np.sign(rng.normal(size=n))
    Z = a[:, None] * rng.normal(size=(n, k))
    X = rng.uniform(0.1, 1.9, n)
    S = rng.uniform(0.1, 2, n)
    Xi = rng.uniform(0.1, 2, n)
    r = rng.normal(size=3 * n + 1)
    step = O.newton_direction(Z, a, C, X, S, Xi, r)
    Q = Z @ Z.T
    J = np.zeros((3 * n + 1, 3 * n + 1))
    I = np.eye(n)
    J[:n, :n] = -Q
    J[:n, n] = a
    J[:n, n + 1:2 * n + 1] = I
    J[:n, 2 * n + 1:] = -I
    J[n, :n] = a
    J[n + 1:2 * n + 1, :n] = np.diag(S)
    J[n + 1:2 * n + 1, n + 1:2 * n + 1] = np.diag(X)
    J[2 * n + 1:, :n] = -np.diag(Xi)
    J[2 * n + 1:, 2 * n + 1:] = np.diag(C - X)
    ref = np.linalg.solve(J, r)
    assert np.allclose(step, ref, rtol=1e-8, atol=1e-8)


def test_write_model_golden(tmp_path):
    m = np.load(os.path.join(G, "model_inputs.npz"))
    p = str(tmp_path / "m.txt")
    O.write_model_rbf(p, m["alpha"], m["labels"], m["inst"], float(m["C"]), float(m["gamma"]),
                      float(m["b"]), int(m["pos"]), int(m["neg"]))
    assert open(p).read() == open(os.path.join(G, "model_rbf.txt")).read()
    p2 = str(tmp_path / "l.txt")
    O.write_model_linear(p2, m["w"], float(m["b"]), int(m["pos"]), int(m["neg"]))
    assert open(p2).read() == open(os.path.join(G, "model_linear.txt")).read()


def test_reader_and_labels_golden():
    g = np.load(os.path.join(G, "tiny_libsvm.npz"))
    lab, x = O.libsvm_read(os.path.join(G, "tiny.libsvm"))
    assert np.array_equal(x, g["x"]) and np.array_equal(lab, g["labels"])
    mapped, pos, neg, bad = O.label_map(np.array([3, 3, 7, 3, 7, 9], dtype=np.float32))
    assert (pos, neg, bad) == (3, 7, 1)
    assert list(mapped[:5]) == [1, 1, -1, 1, -1]


def test_bias_rbf_matches_kkt():
    """At the IPM optimum y (the equality multiplier) equals -b at free SVs."""
    f = np.load(os.path.join(G, "ipm_svc_b.npz"))
    C = float(f["C"])
    X, Xi, S, info = O.mpc(f["Z"], f["y"], C)
    U = f["y"][:, None] * f["Z"]  # unlabelled factor: U U^T = K
    b, cnt, sd = O.bias_rbf(X, f["y"].astype(np.float32), f["Z"], C)
    assert cnt > 0
    # the free-SV average scatters by its reported std around the multiplier
    assert abs(b - float(f["b"])) <= 3 * sd + 1e-4
    assert abs(b + info.y) <= 3 * sd + 1e-4
    assert abs(info.y + float(f["b"])) < 1e-4


def test_grad_matches_central_differences():
    """The FP64 gradient of the dual through the factor (all three projector terms) and of
    the held-out smoothed hinge, against central differences of the same objectives."""
    from datagen import scaled_mixture
    X, y, g0 = scaled_mixture(360, 5, 21)
    X, Xt, y, yt = X[:280], X[280:], y[:280], y[280:]
    n, d, k = X.shape[0], X.shape[1], 14
    rng = np.random.default_rng(4)
    om = O.gaussrand(n * k)
    alpha = rng.uniform(0.0, 4.0, n)
    alpha[rng.random(n) < 0.4] = 0.0
    lg, s = np.log(g0), rng.normal(0.0, 0.3, d)
    D, L, gD, gL, st = O.grad(X, y, k, om, alpha, 0.15, lg, s, Xt, yt)
    assert st == 0
    h = 1e-5
    for m in range(d + 1):
        e = np.zeros(d + 1)
        e[m] = h
        Dp, Lp, _, _, _ = O.grad(X, y, k, om, alpha, 0.15, lg + e[0], s + e[1:], Xt, yt, False)
        Dm, Lm, _, _, _ = O.grad(X, y, k, om, alpha, 0.15, lg - e[0], s - e[1:], Xt, yt, False)
        assert abs((Dp - Dm) / (2 * h) - gD[m]) <= 1e-6 * np.abs(gD).max(), (m, gD[m])
        assert abs((Lp - Lm) / (2 * h) - gL[m]) <= 1e-6 * np.abs(gL).max(), (m, gL[m])
    # log gamma acts like a common shift of the log scales (K depends on gamma e^{2 s_f})
    assert abs(gD[0] - 0.5 * gD[1:].sum()) <= 1e-9 * np.abs(gD).max()


def test_decision_and_linear_bias_against_sklearn():
    """orc_decision and orc_bias_linear on sklearn's own solutions: the exact-kernel decision
    values of an RBF SVC (dual_coef_ = y_j alpha_j over its support vectors, intercept_ = b)
    and the bias of a linear SVC (the mean of y_i - w.x_i over free support vectors, which
    at the optimum is the intercept)."""
    from sklearn.svm import SVC
    rng = np.random.default_rng(17)
    n, d, C, gamma = 300, 5, 2.0, 0.3
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    X = (rng.normal(size=(n, d)) + 0.6 * y[:, None]).astype(np.float32)
    Xt = rng.normal(size=(80, d)).astype(np.float32)
    clf = SVC(kernel="rbf", gamma=gamma, C=C, tol=1e-10).fit(X.astype(np.float64), y)
    # sklearn's convention: f = dual_coef_ . K(sv, x) + intercept_, positive for classes_[1]
    sv = clf.support_
    K = np.exp(-gamma * ((X[sv].astype(np.float64)[:, None, :] - Xt[None].astype(np.float64)) ** 2).sum(-1))
    f_sk = clf.decision_function(Xt.astype(np.float64))
    assert np.allclose(clf.dual_coef_[0] @ K + clf.intercept_[0], f_sk, rtol=1e-10, atol=1e-12)
    assert list(clf.classes_) == [-1.0, 1.0]
    alpha = np.zeros(n)
    alpha[sv] = clf.dual_coef_[0] * y[sv]
    assert np.all(alpha[sv] > 0)
    keep = alpha > 1e-6  # the oracle skips multipliers at or below 1e-6, as the reference
    f_or = O.decision(alpha, y.astype(np.float32), X, gamma, clf.intercept_[0], Xt)
    f_ref = (clf.dual_coef_[0][keep[sv]] @ K[keep[sv]]) + clf.intercept_[0]
    assert np.allclose(f_or, f_ref, rtol=1e-12, atol=1e-12)
    assert np.abs(f_or - f_sk).max() <= 1e-5 * np.abs(f_sk).max()
    lin = SVC(kernel="linear", C=C, tol=1e-12).fit(X.astype(np.float64), y)
    al = np.zeros(n)
    al[lin.support_] = lin.dual_coef_[0] * y[lin.support_]
    b, w = O.bias_linear(al, y.astype(np.float32), X, C)
    assert np.allclose(w, lin.coef_[0], rtol=1e-9, atol=1e-9)
    assert abs(b - lin.intercept_[0]) <= 1e-5 * (1 + abs(lin.intercept_[0])), (b, lin.intercept_)


def _scaled_kernel(Xa, Xb, gamma, s):
    """exp(-gamma sum_f e^{2 s_f} (x_af - x_bf)^2) in FP64."""
    sc = np.exp(s)
    return np.exp(-gamma * cdist(Xa.astype(np.float64) * sc, Xb.astype(np.float64) * sc, "sqeuclidean"))


def test_grad_values_closed_form():
    """The values behind orc_grad, restated with numpy: D = sum alpha - 1/2 alpha^T P K_Y P
    alpha with P = Q Q^T, Q an orthonormal basis of K_Y Omega (numpy QR), and L the mean
    smoothed hinge of the exact-kernel decision values.  A projector onto range(Omega)
    instead of range(K_Y Omega), or an unlabelled kernel, fails it."""
    from datagen import scaled_mixture
    X, y, g0 = scaled_mixture(200, 4, 5)
    X, Xt, y, yt = X[:150], X[150:], y[:150], y[150:]
    n, d, k = X.shape[0], X.shape[1], 10
    rng = np.random.default_rng(8)
    om = O.gaussrand(n * k)
    alpha = rng.uniform(0.0, 2.0, n)
    alpha[rng.random(n) < 0.3] = 0.0
    lg, s, b = np.log(g0), rng.normal(0.0, 0.3, d), -0.2
    D, L, _, _, st = O.grad(X, y, k, om, alpha, b, lg, s, Xt, yt, False)
    assert st == 0
    yd = y.astype(np.float64)
    K = _scaled_kernel(X, X, g0, s)
    Ky = yd[:, None] * K * yd[None, :]
    Q, _ = np.linalg.qr(Ky @ om.reshape(n, k).astype(np.float64))
    P = Q @ Q.T
    D_np = alpha.sum() - 0.5 * alpha @ (P @ Ky @ P) @ alpha
    f = _scaled_kernel(Xt, X, g0, s) @ (alpha * yd) + b
    z = yt.astype(np.float64) * f
    h = np.where(z <= 0, 0.5 - z, np.where(z < 1, 0.5 * (1 - z) ** 2, 0.0))
    assert abs(D - D_np) <= 1e-10 * abs(D_np)
    assert abs(L - h.mean()) <= 1e-12 * max(1.0, abs(L))
    # the plausible mistakes give other values
    Qo, _ = np.linalg.qr(om.reshape(n, k).astype(np.float64))
    Po = Qo @ Qo.T
    assert abs(D - (alpha.sum() - 0.5 * alpha @ (Po @ Ky @ Po) @ alpha)) > 1e-3 * abs(D_np)
    assert abs(D - (alpha.sum() - 0.5 * alpha @ (P @ K @ P) @ alpha)) > 1e-3 * abs(D_np)


def test_grad_dual_is_trained_objective():
    """At the solver's optimum, with no eigenvalue clipped, D(theta_0) is minus the primal
    objective the IPM reports: U U^T = P K_Y P and pobj = 1/2 a^T U U^T a - sum a."""
    from datagen import scaled_mixture
    X, y, g0 = scaled_mixture(400, 5, 9)
    n, d, k = X.shape[0], X.shape[1], 16
    om = O.gaussrand(n * k)
    U, w, realk = O.lra(X, y, k, g0, om.reshape(n, k))
    assert realk == k
    al, xi, s, info = O.mpc(U, y.astype(np.float64), 1.0, stop="gap", tol=1e-9)
    assert info.reason == 0
    D, L, _, _, st = O.grad(X, y, k, om, al, 0.0, np.log(g0), np.zeros(d), want_grad=False)
    assert st == 0
    pobj = 0.5 * al @ (U @ (U.T @ al)) - al.sum()
    assert abs(D + pobj) <= 1e-9 * abs(pobj)
    assert abs(pobj - info.pobj) <= 1e-7 * abs(pobj)


def test_lra_reference_fp32_storage_mode():
    """LRA_REF_FP32 (the reference's float W, Q, C, X and FP32 U = W X): the factor is
    FP32-representable, agrees with the FP64 mode to FP32 rounding, and is exact at k = n to
    FP32 accuracy."""
    from datagen import gaussian_mixture
    X, y = gaussian_mixture(400, 6, 12)
    n, k = 400, 24
    om = O.gaussrand(n * k).reshape(n, k)
    U0, w0, r0 = O.lra(X, y, k, 0.1, om, O.LRA_FP64)
    U1, w1, r1 = O.lra(X, y, k, 0.1, om, O.LRA_REF_FP32)
    assert np.array_equal(U1, U1.astype(np.float32).astype(np.float64))
    G0, G1 = U0 @ U0.T, U1 @ U1.T
    rel = np.abs(G1 - G0).max() / np.abs(G0).max()
    assert 1e-9 < rel < 1e-5, rel  # rounded, but only by FP32 storage
    assert r0 == r1 and np.allclose(w1, w0, rtol=1e-5, atol=1e-6 * w0[-1])
    # k = n: range(K_Y Omega) is everything, U U^T = K_Y to FP32 accuracy
    m = 60
    Xs, ys = X[:m], y[:m]
    Un, _, _ = O.lra(Xs, ys, m, 0.3, O.gaussrand(m * m).reshape(m, m), O.LRA_REF_FP32)
    yd = ys.astype(np.float64)
    Ky = yd[:, None] * np.exp(-0.3 * cdist(Xs.astype(np.float64), Xs.astype(np.float64), "sqeuclidean")) * yd[None, :]
    assert np.abs(Un @ Un.T - Ky).max() <= 1e-5 * np.abs(Ky).max()


def test_qp_sensitivity_to_factor_rounding():
    """How far the dual solution moves when the factor is perturbed at FP32 rounding level:
    a relative 1e-7 perturbation of U moves alpha by ~20-120x as much (relative to C) on the
    c0-shaped problems -- the QP is near-degenerate -- but stays within the 1e-4 C bar.
    The bias, a mean over the first <= 50 free SVs, moves by up to its sampling spread when
    a multiplier crosses the 1e-6 free-SV threshold."""
    from datagen import gaussian_mixture
    X, y = gaussian_mixture(3750, 18, 4000)
    X, y = X[:3000], y[:3000]
    n, k, C = 3000, 32, 32.0
    U, _, _ = O.lra(X, y, k, 0.05, O.gaussrand(n * k).reshape(n, k))
    a0, _, _, i0 = O.mpc(U, y.astype(np.float64), C)
    rng = np.random.default_rng(1)
    Up = U * (1 + 1e-7 * rng.standard_normal(U.shape))
    a1, _, _, i1 = O.mpc(Up, y.astype(np.float64), C)
    da = np.abs(a1 - a0).max() / C
    assert i0.reason == i1.reason == 0
    assert 10 * 1e-7 <= da <= 1e-4, da
    assert abs(i1.pobj - i0.pobj) <= 1e-7 * abs(i0.pobj)
