"""Gradients of the trained model in (log gamma, per-feature log-scales), tsvm_grad_dev, against
the FP64 oracle gradient (itself pinned to central differences on the CPU), against central
differences of the GPU objectives (h = 1e-3), and through a torch.autograd.Function."""
import numpy as np
import pytest

import oracle_lib as O
from datagen import scaled_mixture

pytestmark = pytest.mark.gpu


def _trained(ctx, n, nt, d, k, seed, C=4.0):
    """A model trained by the library at s = 0 (FP32 form): X, y, held-out rows, Omega,
    alpha, b and gamma."""
    import tensorsvm_amd as T
    X, y, g0 = scaled_mixture(n + nt, d, seed)
    X, Xt, y, yt = X[:n], X[n:], y[:n], y[n:]
    om = O.gaussrand(n * k)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=g0, precision=T.PREC_FP32, omega=om,
                            stop=T.capi.STOP_GAP, tol=1e-6, maxit=200)
    return X, y, Xt, yt, om, al, info.b, g0


def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


def _gpu_grad(ctx, dX, dy, dal, b, k, lg, s, dom, dXt, dyt, want=True):
    n, d = dX.shape
    return ctx.grad_dev(dX, dy, n, d, dal, b, k, lg, s, dom, dXt, dyt, dXt.shape[0], want)


def test_grad_matches_fp64_oracle(ctx):
    n, nt, d, k = 1500, 500, 8, 24
    X, y, Xt, yt, om, al, b, g0 = _trained(ctx, n, nt, d, k, 31)
    s = np.random.default_rng(2).normal(0.0, 0.2, d)
    lg = np.log(g0)
    dX, dy, dal, dom, dXt, dyt = _dev(X, y, al, om, Xt, yt)
    D, L, gD, gL, info = _gpu_grad(ctx, dX, dy, dal, b, k, lg, s, dom, dXt, dyt)
    Do, Lo, gDo, gLo, st = O.grad(X, y, k, om, al, b, lg, s, Xt, yt)
    assert st == 0
    assert abs(D - Do) <= 1e-5 * abs(Do)
    assert abs(L - Lo) <= 1e-5 * max(abs(Lo), 1e-3)
    # FP32 kernel products against FP64: relative to the gradient's scale
    assert np.abs(gD - gDo).max() <= 2e-3 * np.abs(gDo).max(), (gD, gDo)
    assert np.abs(gL - gLo).max() <= 2e-3 * np.abs(gLo).max(), (gL, gLo)
    assert abs(gD[0] - gDo[0]) <= 1e-3 * abs(gDo[0])
    assert abs(gL[0] - gLo[0]) <= 1e-3 * abs(gLo[0])
    assert info.backward_flops > 0 and info.t_backward > 0


def test_grad_central_differences(ctx):
    """The log-gamma gradient within 1e-3 relative of central differences (h = 1e-3) of
    the GPU objectives themselves; the per-feature ones within 1e-2 of the gradient scale."""
    n, nt, d, k = 4000, 1000, 12, 48
    X, y, Xt, yt, om, al, b, g0 = _trained(ctx, n, nt, d, k, 77)
    dX, dy, dal, dom, dXt, dyt = _dev(X, y, al, om, Xt, yt)
    lg, s = np.log(g0), np.zeros(d)
    _, _, gD, gL, _ = _gpu_grad(ctx, dX, dy, dal, b, k, lg, s, dom, dXt, dyt)
    h = 1e-3
    for m in range(d + 1):
        e = np.zeros(d + 1)
        e[m] = h
        Dp, Lp, _, _, _ = _gpu_grad(ctx, dX, dy, dal, b, k, lg + e[0], s + e[1:], dom, dXt, dyt, False)
        Dm, Lm, _, _, _ = _gpu_grad(ctx, dX, dy, dal, b, k, lg - e[0], s - e[1:], dom, dXt, dyt, False)
        fD, fL = (Dp - Dm) / (2 * h), (Lp - Lm) / (2 * h)
        if m == 0:
            assert abs(fD - gD[0]) <= 1e-3 * abs(gD[0]), (fD, gD[0])
            assert abs(fL - gL[0]) <= 1e-3 * abs(gL[0]), (fL, gL[0])
        else:
            assert abs(fD - gD[m]) <= 1e-2 * np.abs(gD).max(), (m, fD, gD[m])
            assert abs(fL - gL[m]) <= 1e-2 * np.abs(gL).max(), (m, fL, gL[m])


def test_grad_fp16_forward_close_to_fp32(ctx):
    """The forward pass's factor products in the fp16 product: values and gradients close
    to the FP32 forward's (the backward is FP32 in both)."""
    import tensorsvm_amd as T
    n, nt, d, k = 4000, 1000, 12, 48
    X, y, Xt, yt, om, al, b, g0 = _trained(ctx, n, nt, d, k, 77)
    dX, dy, dal, dom, dXt, dyt = _dev(X, y, al, om, Xt, yt)
    lg, s = np.log(g0), np.zeros(d)
    D32, L32, gD32, gL32, _ = ctx.grad_dev(dX, dy, n, d, dal, b, k, lg, s, dom, dXt, dyt, nt,
                                           precision=T.PREC_FP32)
    D16, L16, gD16, gL16, _ = ctx.grad_dev(dX, dy, n, d, dal, b, k, lg, s, dom, dXt, dyt, nt,
                                           precision=T.PREC_BF16)
    assert abs(D16 - D32) <= 1e-3 * abs(D32), (D16, D32)
    assert L16 == L32  # the held-out loss does not use the factor
    assert np.abs(gD16 - gD32).max() <= 1e-2 * np.abs(gD32).max(), (gD16, gD32)
    assert np.abs(gL16 - gL32).max() <= 1e-6 * np.abs(gL32).max()


def test_grad_autograd_function(ctx):
    """The C-ABI gradient as the backward of a torch.autograd.Function over theta."""
    import torch
    n, nt, d, k = 1200, 400, 6, 16
    X, y, Xt, yt, om, al, b, g0 = _trained(ctx, n, nt, d, k, 5)
    dX, dy, dal, dom, dXt, dyt = _dev(X, y, al, om, Xt, yt)

    class TrainedObjectives(torch.autograd.Function):
        @staticmethod
        def forward(fctx, theta):
            th = theta.detach().cpu().numpy()
            D, L, gD, gL, _ = _gpu_grad(ctx, dX, dy, dal, b, k, th[0], th[1:], dom, dXt, dyt)
            fctx.save_for_backward(torch.from_numpy(gD), torch.from_numpy(gL))
            return torch.tensor([D, L], dtype=torch.float64)

        @staticmethod
        def backward(fctx, gout):
            gD, gL = fctx.saved_tensors
            return gout[0] * gD + gout[1] * gL

    theta = torch.zeros(d + 1, dtype=torch.float64)
    theta[0] = np.log(g0)
    theta.requires_grad_(True)
    out = TrainedObjectives.apply(theta)
    (out[0] + 100.0 * out[1]).backward()
    th = theta.detach().numpy()
    Do, Lo, gDo, gLo, st = O.grad(X, y, k, om, al, b, th[0], th[1:], Xt, yt)
    ref = gDo + 100.0 * gLo
    assert abs(out[0].item() - Do) <= 1e-5 * abs(Do)
    assert np.abs(theta.grad.numpy() - ref).max() <= 2e-3 * np.abs(ref).max()
