"""Weighted Gram of the capacitance, M = I + Z^T diag(w) Z (tsvm_weighted_syrk_dev), against
an FP64 numpy reference: the FP64 form to rounding, the hybrid form (BF16 or FP8 dense
pass, compensated hi + lo bf16 rows, FP64 rows) within the error its split promises."""
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


def _bound(info):
    """Error bound of a hybrid split: 2^-9 per unit of the BF16 rows' sqrt(sum t^2), 2^-16
    per unit of the compensated rows', the FP8 rows' own bound (zero-mean rounding)."""
    return 2.0 ** -9 * info.bf16_bound + 2.0 ** -16 * info.x2_bound + info.fp8_bound


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


@pytest.mark.parametrize("n,k", [(100003, 512), (40000, 1024), (30000, 100), (5000, 64)])
def test_hybrid_within_bound(ctx, n, k):
    import tensorsvm_amd as T
    from conftest import parity_log
    Z, w, M = _case(n, k, 7 * n + k)
    G, info = _run(ctx, Z, w, T.PREC_BF16)
    assert 0 < info.n_fp64_rows < n  # a real split at these weights
    if k > 128:  # the compensated tier takes a middle band (kpt >= 256)
        assert info.n_x2_rows > 0 and info.n_fp64_rows + info.n_x2_rows < n
        assert info.x2_bound <= 4.0 * 128
    else:
        assert info.n_x2_rows == 0
    err = np.linalg.norm(G - M, 2)
    parity_log(f"syrk_hybrid_{n}_{k}", err=err, bound=_bound(info), fp64_rows=int(info.n_fp64_rows),
               x2_rows=int(info.n_x2_rows))
    # three bf16 roundings per term (2^-9 each), zero-mean: the norm of the error sits far
    # below 2^-9 times the bound on sqrt(sum t^2) over the BF16 rows (2^-16 for the
    # compensated rows)
    assert info.bf16_bound <= 4.0
    assert err <= _bound(info) + 1e-12 * np.abs(M).max(), (err, info.as_dict())
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
    assert info.n_fp64_rows == 0 and info.n_x2_rows == 0
    tn = w * (Z.astype(np.float64) ** 2).sum(1)
    assert np.linalg.norm(G - M, 2) <= 2.0 ** -9 * 3 * np.sqrt((tn ** 2).sum()) * 4


def test_hybrid_deterministic(ctx):
    import tensorsvm_amd as T
    Z, w, _ = _case(50001, 256, 11)
    G1, i1 = _run(ctx, Z, w, T.PREC_BF16)
    G2, i2 = _run(ctx, Z, w, T.PREC_BF16)
    assert i1.n_fp64_rows == i2.n_fp64_rows
    assert np.array_equal(G1, G2)


def _fp8_case(n, k, seed, heavy=0.03):
    """Late-solve weights: most rows tiny (non-support vectors, t ~ 1e-8..1e-4), a few
    percent heavy (free support vectors, t ~ 1..100): the FP8 pass's regime."""
    rng = np.random.default_rng(seed)
    Z = (rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)
    w = 10.0 ** rng.uniform(-8.0, -4.0, n)
    hv = rng.random(n) < heavy
    w[hv] = 10.0 ** rng.uniform(0.0, 2.0, int(hv.sum()))
    Zd = Z.astype(np.float64)
    M = np.eye(k) + Zd.T @ (w[:, None] * Zd)
    return Z, w, M


@pytest.mark.parametrize("n,k", [(100003, 1024), (40000, 1024), (30011, 900)])
def test_fp8_pass_within_bound(ctx, n, k):
    """TSVM_PREC_FP8 with late-solve weights at widths where the FP8 pass pays (kpt = 1024):
    the small-t rows go through the e4m3 pass (per-column power-of-two scales), the rest
    through FP64, none through BF16; the error of M in the 2-norm is within the bound the
    split reports."""
    import tensorsvm_amd as T
    from conftest import parity_log
    Z, w, M = _fp8_case(n, k, 13 * n + k)
    G, info = _run(ctx, Z, w, T.PREC_FP8)
    assert info.n_fp8_rows >= 0.9 * n, info.as_dict()
    assert info.n_fp8_rows + info.n_fp64_rows + info.n_x2_rows == n
    assert info.bf16_bound == 0.0
    assert 0.0 < info.fp8_bound <= 2.0 ** -5
    err = np.linalg.norm(G - M, 2)
    G16, i16 = _run(ctx, Z, w, T.PREC_BF16)
    err16 = np.linalg.norm(G16 - M, 2)
    parity_log(f"syrk_fp8_{n}_{k}", err_fp8=err, bound_fp8=info.fp8_bound, err_bf16=err16,
               bound_bf16=2.0 ** -9 * i16.bf16_bound, fp8_rows=int(info.n_fp8_rows),
               fp64_rows=int(info.n_fp64_rows))
    assert err <= _bound(info) + 1e-12 * np.abs(M).max(), (err, info.as_dict())
    assert np.array_equal(G, G.T)
    # the FP8 pass really ran: its error is far above the FP64 part's rounding
    assert err > 1e-9


def test_fp8_form_falls_back_to_bf16(ctx):
    """The FP8 form keeps the BF16 split, bit-identical to the BF16 form, where the FP8
    pass would not pay: weights spread over ten decades (too many rows between the FP8
    and FP64 tiers), any weights at kpt <= 512 (packing outweighs the faster SYRK), and
    k <= 128 (no FP8 pass at all)."""
    import tensorsvm_amd as T
    for n, k, late in [(100003, 1024, False), (100003, 512, True), (20000, 300, True),
                       (30000, 100, True)]:
        Z, w, _ = (_fp8_case if late else _case)(n, k, 7 * n + k)
        G8, i8 = _run(ctx, Z, w, T.PREC_FP8)
        G16, i16 = _run(ctx, Z, w, T.PREC_BF16)
        assert i8.n_fp8_rows == 0 and i8.fp8_bound == 0.0
        assert i8.n_fp64_rows == i16.n_fp64_rows
        assert np.array_equal(G8, G16)


def test_fp8_pass_all_rows(ctx):
    """Every row tiny: all rows in the FP8 pass, no FP64 rows; the relative error of
    Z^T diag(w) Z sits at the e4m3 level (a few 2^-4 per term, averaged over rows)."""
    import tensorsvm_amd as T
    n, k = 50000, 1024
    rng = np.random.default_rng(21)
    Z = (rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)
    w = 10.0 ** rng.uniform(-7.0, -6.0, n)
    Zd = Z.astype(np.float64)
    S = Zd.T @ (w[:, None] * Zd)
    G, info = _run(ctx, Z, w, T.PREC_FP8)
    assert info.n_fp8_rows == n and info.n_fp64_rows == 0 and info.n_x2_rows == 0
    E = G - np.eye(k) - S
    rel = np.linalg.norm(E, 2) / np.linalg.norm(S, 2)
    assert np.linalg.norm(E, 2) <= info.fp8_bound
    assert 1e-5 < rel < 2.0 ** -4, rel


@pytest.mark.parametrize("n,k", [(20000, 512), (12345, 300), (9001, 299)])
def test_compensated_tier_all_rows(ctx, n, k):
    """Weights with t_i ~ 1..10 everywhere: too large for the BF16 pass to take more than a
    few rows, small enough for the compensated tier to take them all, so no dense pass runs
    (the cost model's layout C) and no row is FP64.  The error is at the hi + lo level:
    within 2^-16 of the tier's sqrt(sum t^2) and some 2^-18 of |Z^T diag(w) Z| -- two orders
    below a single bf16 rounding."""
    import tensorsvm_amd as T
    from conftest import parity_log
    rng = np.random.default_rng(n + k)
    Z = (rng.standard_normal((n, k)) / np.sqrt(k)).astype(np.float32)
    w = 10.0 ** rng.uniform(0.0, 1.0, n)
    Zd = Z.astype(np.float64)
    S = Zd.T @ (w[:, None] * Zd)
    G, info = _run(ctx, Z, w, T.PREC_BF16, budget=16.0)
    assert info.n_x2_rows == n and info.n_fp64_rows == 0, info.as_dict()
    assert info.bf16_bound == 0.0
    E = G - np.eye(k) - S
    err = np.linalg.norm(E, 2)
    rel = err / np.linalg.norm(S, 2)
    parity_log(f"syrk_x2_{n}_{k}", err=err, bound=_bound(info), rel=rel)
    assert err <= _bound(info)
    assert rel < 2.0 ** -14, rel
    assert np.array_equal(G, G.T)
