"""The CPU baseline's BLAS restatement (oracle/cpu_blas.py) against the C oracle: the tiled
FP32 kernel product and the FP64 interior-point solver give the oracle's results."""
import os
import sys

import numpy as np
import torch

import oracle_lib as O

sys.path.insert(0, os.path.join(O.ROOT, "oracle"))
import cpu_blas  # noqa: E402


def test_tiled_kernel_product_matches_oracle():
    rng = np.random.default_rng(0)
    m, n, d, r = 150, 230, 7, 9
    X1 = rng.normal(size=(m, d)).astype(np.float32)
    X2 = rng.normal(size=(n, d)).astype(np.float32)
    y1 = np.sign(rng.normal(size=m)).astype(np.float32)
    y2 = np.sign(rng.normal(size=n)).astype(np.float32)
    A = rng.normal(size=(n, r)).astype(np.float32)
    W = cpu_blas.kernel_product_tiles(*(torch.from_numpy(v) for v in (X1, y1, X2, y2, A)), 0.2,
                                      B=64).numpy()
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.2)
    assert np.abs(W - ref).max() <= 2e-5 * np.abs(ref).max()


def test_blas_mpc_matches_oracle():
    for name in ("a", "b"):
        f = np.load(os.path.join(O.ROOT, "tests", "golden", f"ipm_svc_{name}.npz"))
        C = float(f["C"])
        Z, y = f["Z"].astype(np.float64), f["y"].astype(np.float64)
        Xo, Xio, So, io = O.mpc(Z, y, C)
        x, xi, s, info = cpu_blas.mpc(torch.from_numpy(Z), torch.from_numpy(y), C)
        assert info["reason"] == io.reason == 0
        assert info["iters"] == io.iters
        assert np.abs(x.numpy() - Xo).max() <= 1e-8 * C
        assert abs(info["pobj"] - io.pobj) <= 1e-10 * (1 + abs(io.pobj))


def test_blas_mpc_iteration_limit_and_gap_rule():
    rng = np.random.default_rng(4)
    n, k, C = 800, 12, 4.0
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    Z = y[:, None] * (rng.normal(size=(n, k)) * 0.6 + 0.1 * y[:, None])
    Xo, _, _, io = O.mpc(Z, y, C, stop="gap", tol=1e-6)
    x, _, _, info = cpu_blas.mpc(torch.from_numpy(Z), torch.from_numpy(y), C, stop="gap", tol=1e-6)
    assert info["iters"] == io.iters and info["reason"] == io.reason
    assert np.abs(x.numpy() - Xo).max() <= 1e-8 * C
    Xo3, _, _, io3 = O.mpc(Z, y, C, maxit=3)
    x3, _, _, info3 = cpu_blas.mpc(torch.from_numpy(Z), torch.from_numpy(y), C, iter_limit=3)
    assert info3["iters"] == 3 and io3.iters == 3
    assert np.abs(x3.numpy() - Xo3).max() <= 1e-10 * C
