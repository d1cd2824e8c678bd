"""TEST INFRASTRUCTURE ONLY: the reference's plain method on the host cores through a CPU
BLAS/LAPACK (torch's CPU kernels, MKL), used as the CPU baseline of bench.py and checked
against the C oracle in tests/.  Nothing in the product (tensorsvm_amd/) imports it.

* kernel_product_tiles: rbf_kermatmul (tensorsvm-train.cu:2476-2567) -- per 8192 x 8192
  tile the norms (vecnorm 2569-2581), XIJ = X_i X_j^T (the XIJ sgemm, 2506-2518), the
  labelled kernel tile y_i y_j exp(-gamma (XI + XJ - 2 XIJ)) (rbf_kergen 2598-2612) and
  W_i += tile A_j (the K A sgemm with beta = 1, 2537-2555), all in FP32.
* mpc: the interior-point solver (mpc 1507-1785, NewtonStep 1794-1822, SMWSolve
  1824-1840) in FP64: the capacitance I + Zs^T Zs with Zs = D^{-1/2} Z as one GEMM (the
  dsyrk of 1660-1706), its Cholesky (dpotrf 1709), the SMW solves as two matrix-vector
  products and a triangular solve pair (dpotrs).
"""
import math

import torch


def kernel_product_tiles(X1, y1, X2, y2, A, gamma, B=8192):
    """W = diag(y1) K(X1, X2) diag(y2) A in FP32 over B x B tiles (torch CPU tensors)."""
    X1, X2 = X1.float(), X2.float()
    y1, y2, A = y1.float(), y2.float(), A.float()
    m, n = X1.shape[0], X2.shape[0]
    W = torch.zeros(m, A.shape[1], dtype=torch.float32)
    for i in range(0, m, B):
        xi = X1[i:i + B]
        XI = (xi * xi).sum(1)
        for j in range(0, n, B):
            xj = X2[j:j + B]
            XJ = (xj * xj).sum(1)
            XIJ = xi @ xj.T
            buf = torch.exp(-gamma * (XI[:, None] + XJ[None, :] - 2.0 * XIJ))
            buf *= y1[i:i + B, None] * y2[None, j:j + B]
            W[i:i + B] += buf @ A[j:j + B]
    return W


def mpc(Z, a, C, maxit=100, stop="ref", tol=1e-5, iter_limit=None):
    """Mehrotra predictor-corrector on Q = Z Z^T (Z n x k, a labels), FP64 on the CPU.
    Returns (x, xi, s, info dict).  iter_limit ends the loop early (timing samples) without
    changing the iterates computed up to it."""
    Z = Z.double()
    a = a.double()
    n, k = Z.shape
    dev = dict(dtype=torch.float64)
    one = torch.ones(n, **dev)
    e = Z @ (Z.T @ one)                          # Q 1 (1535-1552)
    ox = float((e.sum() / (e * e).sum()).item())
    ox = min(max(ox, 0.01 * C), 0.99 * C)
    x = torch.full((n,), ox, **dev)
    s = torch.ones(n, **dev)
    xi = torch.ones(n, **dev)
    y = 0.0
    info = {"reason": 2}

    def smw(L, D, v):                            # SMWSolve (1824-1840)
        v = v / D
        t = torch.cholesky_solve((Z.T @ v)[:, None], L)[:, 0]
        return v - (Z @ t) / D

    def newton(L, D, r1, r2, r3, r4):            # NewtonStep (1794-1822)
        r5 = r1 - r3 / x + r4 / (C - x)
        r7 = smw(L, D, r5)
        b = smw(L, D, a)
        dy = (r2 + float(a @ r7)) / float(a @ b)
        dx = smw(L, D, a * dy - r5)
        return dx, dy, (r3 - s * dx) / x, (r4 + xi * dx) / (C - x)

    def step(dx, ds, dxi):                       # 1733-1740
        al = 1.0
        neg, pos = dx < 0, dx > 0
        if neg.any():
            al = min(al, float((-x[neg] / dx[neg]).min()))
        if pos.any():
            al = min(al, float(((C - x[pos]) / dx[pos]).min()))
        if (ds < 0).any():
            al = min(al, float((-s[ds < 0] / ds[ds < 0]).min()))
        if (dxi < 0).any():
            al = min(al, float((-xi[dxi < 0] / dxi[dxi < 0]).min()))
        return al

    it = 0
    for it in range(maxit):
        if iter_limit is not None and it >= iter_limit:
            break
        mu = float((x * s + (C - x) * xi).sum()) / (2 * n)
        e = Z @ (Z.T @ x)                        # 1616-1619
        r1 = e - a * y - s + xi - 1.0
        r2 = -float(a @ x)
        r3 = -s * x
        r4 = -xi * (C - x)
        D = s / x + xi / (C - x)                 # 1624-1630
        cond = float(D.max() / D.min())
        normdx, normdy = float(r1.norm()), abs(r2)
        tmp = 0.5 * float(x @ e)                 # 1636-1643
        pobj, dobj = tmp - float(x.sum()), -tmp - C * float(xi.sum())
        gap = (pobj - dobj) / (1 + abs(pobj))
        info = {"iters": it, "mu": mu, "normdx": normdx, "normdy": normdy, "pobj": pobj,
                "dobj": dobj, "gap_rel": gap, "condD": cond, "y": y, "reason": 2}
        if stop == "ref":                        # 1647
            conv = mu < 1e-7 and normdx < 1e-7 and normdy < 1e-7
        else:
            conv = (gap <= tol and normdx / (1 + math.sqrt(n)) <= tol
                    and normdy <= tol * (1 + float(x.abs().sum()) / C))
        if conv:
            info["reason"] = 0
            break
        if cond > 1e18:                          # 1653
            info["reason"] = 1
            break
        Zs = Z / D.sqrt()[:, None]               # 1660-1706
        M = Zs.T @ Zs
        M.diagonal().add_(1.0)
        L, fail = torch.linalg.cholesky_ex(M)    # 1709
        if int(fail) != 0:
            info["reason"] = 3
            break
        dx, dy, ds, dxi = newton(L, D, r1, r2, r3, r4)          # predictor (1727)
        al = step(dx, ds, dxi)
        mu_aff = float(((x + al * dx) * (s + al * ds)
                        + (C - x - al * dx) * (xi + al * dxi)).sum()) / (2 * n)
        sigma = (mu_aff / mu) ** 3
        c3 = sigma * mu - dx * ds                               # corrector (1750-1758)
        c4 = sigma * mu + dx * dxi
        cx, cy, cs, cxi = newton(L, D, torch.zeros(n, **dev), 0.0, c3, c4)
        dx, dy, ds, dxi = dx + cx, dy + cy, ds + cs, dxi + cxi
        al = 0.99 * step(dx, ds, dxi)                           # 1769-1777
        x = x + al * dx
        s = s + al * ds
        xi = xi + al * dxi
        y += al * dy
    else:
        it = maxit
    info["iters"] = it
    return x, xi, s, info
