"""The benchmark configurations' shapes at sizes the oracle finishes in seconds:
c1 (covtype-like, d = 54, gamma = 1, k = 400: padded factor width 416), c3 (the
differentiation form, d = 32, k = 256) and c4 (d = 18, k = 1024: the widest factor, whose
residual pass runs unfused).  Each runs the HIP path through the C-ABI against the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import parity_log
from datagen import covtype_like, gaussian_mixture, scaled_mixture
from test_gpu_train import check_train_parity

pytestmark = pytest.mark.gpu


def _c1_data():
    X, y = covtype_like(6250, 1101)
    return X[:5000], y[:5000], X[5000:], y[5000:]


def test_c1_fp32_train_matches_oracle(ctx):
    """c1 shape: the plain FP32 form end to end against the oracle's plain method."""
    import tensorsvm_amd as T
    X, y, Xt, yt = _c1_data()
    k, gamma, C = 400, 1.0, 32.0
    om = O.gaussrand(X.shape[0] * k)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=gamma, precision=T.PREC_FP32, omega=om)
    assert info.ipm.reason == 0
    assert info.lra.realk == k
    check_train_parity("c1_fp32", ctx, X, y, Xt, yt, k, gamma, C, om, al, info)


def test_c1_bf16_train_close_to_fp32(ctx):
    """c1 shape in the BF16 form (fp16 factor products at KD = 176, hybrid capacitance):
    the same optimum as the FP32 form to its rounding, same held-out decisions."""
    import tensorsvm_amd as T
    X, y, Xt, yt = _c1_data()
    k, gamma, C = 400, 1.0, 32.0
    om = O.gaussrand(X.shape[0] * k)
    kw = dict(k=k, C=C, gamma=gamma, omega=om, stop=T.capi.STOP_GAP, tol=1e-5)
    a32, _, i32 = ctx.train(X, y, precision=T.PREC_FP32, **kw)
    a16, _, i16 = ctx.train(X, y, precision=T.PREC_BF16, **kw)
    assert i32.ipm.reason == 0 and i16.ipm.reason == 0
    f32 = ctx.decision_function(X, y, a32, gamma, i32.b, Xt)
    f16 = ctx.decision_function(X, y, a16, gamma, i16.b, Xt)
    acc32, acc16 = np.mean(np.sign(f32) == yt), np.mean(np.sign(f16) == yt)
    agree = np.mean(np.sign(f32) == np.sign(f16))
    parity_log("c1_bf16_vs_fp32", dpobj_rel=abs(i16.ipm.pobj - i32.ipm.pobj) / abs(i32.ipm.pobj),
               acc_fp32=acc32, acc_bf16=acc16, sign_agreement=agree,
               iters_fp32=i32.ipm.iters, iters_bf16=i16.ipm.iters)
    assert abs(i16.ipm.pobj - i32.ipm.pobj) <= 5e-3 * abs(i32.ipm.pobj)
    assert agree >= 0.98, agree
    assert abs(acc16 - acc32) <= 0.01, (acc16, acc32)


def test_c3_gradient_matches_fp64_oracle_and_fd(ctx):
    """c3 shape (d = 32, k = 256): the backward kernel's 3(2d+1) = 195-column right-hand
    side against the FP64 oracle gradient, and d/dlog gamma against central differences
    (h = 1e-3) of the GPU objectives."""
    import torch
    import tensorsvm_amd as T
    n, nt, d, k, C = 3000, 1000, 32, 256, 1.0
    X, y, g0 = scaled_mixture(n + nt, d, 1103)
    X, Xt, y, yt = X[:n], X[n:], y[:n], y[n:]
    om = O.gaussrand(n * k)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=g0, precision=T.PREC_FP32, omega=om,
                            stop=T.capi.STOP_GAP, tol=1e-6, maxit=200)
    assert info.ipm.reason == 0
    dX, dy, dal, dom, dXt, dyt = (torch.from_numpy(np.ascontiguousarray(a)).cuda()
                                  for a in (X, y, al, om, Xt, yt))
    lg, s = np.log(g0), np.zeros(d)
    D, L, gD, gL, gi = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg, s, dom, dXt, dyt, nt)
    Do, Lo, gDo, gLo, st = O.grad(X, y, k, om, al, info.b, lg, s, Xt, yt)
    assert st == 0
    h = 1e-3
    Dp, Lp, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg + h, s, dom, dXt, dyt, nt, False)
    Dm, Lm, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg - h, s, dom, dXt, dyt, nt, False)
    fD, fL = (Dp - Dm) / (2 * h), (Lp - Lm) / (2 * h)
    parity_log("c3_grad", D_rel=abs(D - Do) / abs(Do), L_rel=abs(L - Lo) / abs(Lo),
               gD_rel=np.abs(gD - gDo).max() / np.abs(gDo).max(),
               gL_rel=np.abs(gL - gLo).max() / np.abs(gLo).max(),
               gD0_rel=abs(gD[0] - gDo[0]) / abs(gDo[0]), gL0_rel=abs(gL[0] - gLo[0]) / abs(gLo[0]),
               fdD_rel=abs(fD - gD[0]) / abs(gD[0]), fdL_rel=abs(fL - gL[0]) / abs(gL[0]),
               backward_tflops=gi.backward_tflops)
    assert abs(D - Do) <= 1e-5 * abs(Do)
    assert abs(L - Lo) <= 1e-5 * max(abs(Lo), 1e-3)
    assert np.abs(gD - gDo).max() <= 2e-3 * np.abs(gDo).max()
    assert np.abs(gL - gLo).max() <= 2e-3 * np.abs(gLo).max()
    assert abs(gD[0] - gDo[0]) <= 1e-3 * abs(gDo[0])
    assert abs(gL[0] - gLo[0]) <= 1e-3 * abs(gLo[0])
    assert abs(fD - gD[0]) <= 1e-3 * abs(gD[0]), (fD, gD[0])
    assert abs(fL - gL[0]) <= 1e-3 * abs(gL[0]), (fL, gL[0])


def _projected_kernel(X, y, gamma, om):
    """(P K_Y P)+ with P the projector onto range(K_Y Omega): the factor's U U^T, formed
    from the oracle's FP64 kernel products and numpy's QR and eigh."""
    n, k = om.shape
    W = O.kernel_product(X, y, X, y, om.astype(np.float64), gamma)
    Q, _ = np.linalg.qr(W)
    Cm = Q.T @ O.kernel_product(X, y, X, y, Q, gamma)
    w, V = np.linalg.eigh(0.5 * (Cm + Cm.T))
    F = Q @ (V * np.sqrt(np.maximum(w, 0.0)))
    return F, w


@pytest.fixture(scope="module")
def c4_case():
    X, y = gaussian_mixture(5000, 18, 1104)
    X, y = X[:4000], y[:4000]
    k = 1024
    om = O.gaussrand(X.shape[0] * k).reshape(-1, k)
    F, w = _projected_kernel(X, y, 0.05, om)
    return X, y, om, F, w


@pytest.mark.parametrize("form", ["fp32", "bf16"])
def test_c4_factor_k1024(ctx, c4_case, form):
    """c4 shape (k = 1024): the factor against the FP64 projected kernel."""
    import tensorsvm_amd as T
    X, y, om, F, w = c4_case
    prec = T.PREC_FP32 if form == "fp32" else T.PREC_BF16
    U, info = ctx.lra(X, y, 1024, 0.05, prec, omega=om)
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    Go = F @ F.T
    rel = np.linalg.norm(G - Go) / np.linalg.norm(Go)
    parity_log(f"c4_factor_{form}", rel_diff=rel, realk=info.realk, eig_max_rel=abs(info.eig_max - w[-1]) / w[-1])
    assert rel < (1e-4 if form == "fp32" else 2e-3), rel
    assert np.isclose(info.eig_max, w[-1], rtol=1e-4 if form == "fp32" else 1e-3)


# refine = 2: two sweeps for each of the affine and the corrector's steps; -1 (the default):
# one sweep for the combined step
@pytest.mark.parametrize("refine", [2, -1])
def test_c4_bf16_ipm_k1024_matches_oracle(ctx, c4_case, refine):
    """c4 shape: the BF16 interior-point form (hybrid capacitance at k = 1024, refinement)
    against the oracle's FP64 IPM on the same factor."""
    import tensorsvm_amd as T
    X, y, om, F, w = c4_case
    U, info = ctx.lra(X, y, 1024, 0.05, T.PREC_BF16, omega=om)
    C = 32.0
    al, xi, s, ii = ctx.ipm(U, y, C, precision=T.PREC_BF16, refine=refine, stop=T.capi.STOP_GAP,
                            tol=1e-7)
    Xo, Xio, So, io = O.mpc(U.astype(np.float64), y.astype(np.float64), C, stop="gap", tol=1e-7)
    parity_log(f"c4_ipm_bf16_refine{refine}", dalpha_over_C=np.abs(al - Xo).max() / C,
               dpobj_rel=abs(ii.pobj - io.pobj) / (1 + abs(io.pobj)), iters_gpu=ii.iters,
               iters_oracle=io.iters, bf16_row_frac=ii.bf16_row_frac)
    assert ii.reason == 0 and io.reason == 0
    assert ii.syrk_bf16_iters > 0
    assert abs(ii.pobj - io.pobj) <= 1e-6 * (1 + abs(io.pobj))
    assert np.abs(al - Xo).max() <= 1e-4 * C


def test_c4_fp8_ipm_k1024_matches_oracle(ctx, c4_case):
    """c4 shape in the FP8 form: the iterations whose small-t rows fit the e4m3 budget (and
    whose remaining bulk is small) form the capacitance from the FP8 pass plus FP64 rows;
    the refined Newton steps reach the oracle's FP64 optimum to the BF16 form's bounds."""
    import tensorsvm_amd as T
    X, y, om, F, w = c4_case
    U, info = ctx.lra(X, y, 1024, 0.05, T.PREC_BF16, omega=om)
    C = 32.0
    al, xi, s, ii = ctx.ipm(U, y, C, precision=T.PREC_FP8, refine=2, stop=T.capi.STOP_GAP, tol=1e-7)
    Xo, Xio, So, io = O.mpc(U.astype(np.float64), y.astype(np.float64), C, stop="gap", tol=1e-7)
    parity_log("c4_ipm_fp8", dalpha_over_C=np.abs(al - Xo).max() / C,
               dpobj_rel=abs(ii.pobj - io.pobj) / (1 + abs(io.pobj)), iters_gpu=ii.iters,
               iters_oracle=io.iters, bf16_row_frac=ii.bf16_row_frac, fp8_row_frac=ii.fp8_row_frac)
    assert ii.reason == 0 and io.reason == 0
    assert ii.fp8_row_frac > 0.0  # the FP8 pass ran in some iterations
    assert abs(ii.pobj - io.pobj) <= 1e-6 * (1 + abs(io.pobj))
    assert np.abs(al - Xo).max() <= 1e-4 * C
