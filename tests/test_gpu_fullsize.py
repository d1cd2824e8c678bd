"""Properties that do not depend on size, checked at the full default workload (1.6M
training rows of the 18-feature Gaussian mixture, k = 512): where the oracle cannot run,
the kernel products must still be linear and symmetric, the factor's columns orthogonal
(U = W V sqrt(Lambda) with W^T W = L L^T exactly as used), and the IPM's answer feasible
and converged."""
import numpy as np
import pytest

from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu

N, D, K, GAMMA, C = 1_600_000, 18, 512, 0.05, 32.0


@pytest.fixture(scope="module")
def data():
    import torch
    X, y = gaussian_mixture(2_000_000, D, 1002)
    perm = np.random.default_rng(1003).permutation(2_000_000)[:N]
    X, y = np.ascontiguousarray(X[perm]), np.ascontiguousarray(y[perm])
    dX, dy = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
    return dX, dy


@pytest.mark.parametrize("form", ["fp32", "fp16"])
def test_kernel_product_linear_and_symmetric(ctx, data, form):
    import torch
    import tensorsvm_amd as T
    prec = T.PREC_FP32 if form == "fp32" else T.PREC_FP16
    dX, dy = data
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    r = 8
    A = torch.randn(N, r, device="cuda", generator=g)
    B = torch.randn(N, r, device="cuda", generator=g)
    out = {}
    for name, M in (("A", A), ("B", B), ("AB", 0.75 * A - 1.25 * B)):
        W = torch.empty(N, r, device="cuda")
        ctx.kernel_product_dev(dX, dy, N, dX, dy, N, D, M.contiguous(), r, GAMMA, prec, W)
        out[name] = W.double()
    # linearity: K (0.75 A - 1.25 B) = 0.75 K A - 1.25 K B
    lin = out["AB"] - (0.75 * out["A"] - 1.25 * out["B"])
    tol = 1e-5 if form == "fp32" else 3e-3
    assert lin.norm().item() <= tol * out["AB"].norm().item(), lin.norm().item()
    # symmetry of K_Y: A^T (K_Y B) = (K_Y A)^T B
    s1 = A.double().T @ out["B"]
    s2 = out["A"].T @ B.double()
    assert (s1 - s2).norm().item() <= tol * s1.norm().item()


def test_factor_columns_orthogonal_and_ipm_converged(ctx, data):
    import torch
    import tensorsvm_amd as T
    dX, dy = data
    om = torch.from_numpy(T.gaussrand_fill(N * K)).cuda()
    U = torch.empty(N, K, device="cuda")
    info = ctx.lra_dev(dX, dy, N, D, K, GAMMA, T.PREC_BF16, om, U)
    assert info.realk == K
    G = U.double().T @ U.double()
    lam = torch.diagonal(G)
    # the eigenvalues of the projected kernel, in order, and orthogonal columns
    assert abs(lam.max().item() - info.eig_max) <= 1e-4 * info.eig_max
    assert abs(lam.min().item() - info.eig_min) <= 1e-3 * info.eig_min
    off = G / torch.sqrt(lam[:, None] * lam[None, :]) - torch.eye(K, dtype=torch.float64, device="cuda")
    assert off.abs().max().item() <= 1e-4, off.abs().max().item()
    # the dual IPM on that factor: feasible, converged to the requested gap
    al = torch.empty(N, dtype=torch.float64, device="cuda")
    xi = torch.empty(N, dtype=torch.float64, device="cuda")
    ii = ctx.ipm_dev(U, dy, N, K, C, al, xi, maxit=300, stop=T.capi.STOP_GAP, tol=1e-5,
                     precision=T.PREC_BF16)
    assert ii.reason == 0 and ii.gap_rel <= 1e-5
    a = al.cpu().numpy()
    assert a.min() >= 0.0 and a.max() <= C
    yy = dy.double().cpu().numpy()
    assert abs(np.dot(yy, a)) <= 1e-5 * (1.0 + np.abs(a).sum() / C)
