"""Weighted Gram of the capacitance, M = I + Z^T diag(w) Z (tsvm_weighted_syrk_dev), against
an FP64 numpy reference: the FP64 form to rounding, the hybrid BF16/FP64 form within the
error its split promises."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, k, seed, wlo=-8.0, whi=2.0):
    rng = np.random.default_rng(seed)
    Z = (rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)
    # weights spanning ten orders of magnitude, as D^-1 does late in a solve
    w = 10.0 ** rng.uniform(wlo, whi, n)
    Zd = Z.astype(np.float64)
    M = np.eye(k) + Zd.T @ (w[:, None] * Zd)
    return Z, w, M


def _run(ctx, Z, w, precision, budget=0.0):
    import torch
    n, k = Z.shape
    dZ = torch.from_numpy(Z).cuda()
    dw = torch.from_numpy(w).cuda()
    dM = torch.empty((k, k), dtype=torch.float64, device="cuda")
    info = ctx.weighted_syrk_dev(dZ, n, k, dw, precision, dM, budget)
    return dM.cpu().numpy(), info


@pytest.mark.parametrize("n,k", [(20000, 100), (4097, 512), (777, 33)])
def test_fp64_form(ctx, n, k):
    import tensorsvm_amd as T
    Z, w, M = _case(n, k, n + k)
    G, info = _run(ctx, Z, w, T.PREC_FP32)
    assert info.n_fp64_rows == n
    assert np.abs(G - M).max() <= 1e-12 * np.abs(M).max()
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("n,k", [(100003, 512), (30000, 100), (5000, 64)])
def test_hybrid_within_bound(ctx, n, k):
    import tensorsvm_amd as T
    Z, w, M = _case(n, k, 7 * n + k)
    G, info = _run(ctx, Z, w, T.PREC_BF16)
    assert 0 < info.n_fp64_rows < n  # a real split at these weights
    err = np.linalg.norm(G - M, 2)
    # three bf16 roundings per term (2^-9 each), zero-mean: the norm of the error sits far
    # below 2^-9 times the bound on sqrt(sum t^2) over the BF16 rows
    assert info.bf16_bound <= 4.0
    assert err <= 2.0 ** -9 * info.bf16_bound + 1e-12 * np.abs(M).max(), (err, info.as_dict())
    # the heavy rows are exact: the error is tiny against the matrix itself
    assert err <= 1e-4 * np.linalg.norm(M, 2)
    assert np.array_equal(G, G.T)


def test_hybrid_budget_extremes(ctx):
    import tensorsvm_amd as T
    Z, w, M = _case(20000, 128, 3)
    G64, _ = _run(ctx, Z, w, T.PREC_FP32)
    # a vanishing budget: every row FP64, bit-identical to the FP64 form
    G, info = _run(ctx, Z, w, T.PREC_BF16, budget=1e-300)
    assert info.n_fp64_rows == Z.shape[0]
    assert np.array_equal(G, G64)
    # an unlimited budget: every row BF16, error at bf16 level of the whole matrix
    G, info = _run(ctx, Z, w, T.PREC_BF16, budget=1e300)
    assert info.n_fp64_rows == 0
    tn = w * (Z.astype(np.float64) ** 2).sum(1)
    assert np.linalg.norm(G - M, 2) <= 2.0 ** -9 * 3 * np.sqrt((tn ** 2).sum()) * 4


def test_hybrid_deterministic(ctx):
    import tensorsvm_amd as T
    Z, w, _ = _case(50001, 256, 11)
    G1, i1 = _run(ctx, Z, w, T.PREC_BF16)
    G2, i2 = _run(ctx, Z, w, T.PREC_BF16)
    assert i1.n_fp64_rows == i2.n_fp64_rows
    assert np.array_equal(G1, G2)
