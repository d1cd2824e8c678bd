"""Fused RBF kernel product (HIP, through the C-ABI) against the FP64 oracle."""
import numpy as np
import pytest

import oracle_lib as O
from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu


def _case(m, n, d, r, gamma, seed):
    rng = np.random.default_rng(seed)
    X1 = rng.standard_normal((m, d)).astype(np.float32) * 0.7
    X2 = rng.standard_normal((n, d)).astype(np.float32) * 0.7
    y1 = np.where(rng.random(m) < 0.5, 1.0, -1.0).astype(np.float32)
    y2 = np.where(rng.random(n) < 0.5, 1.0, -1.0).astype(np.float32)
    A = rng.standard_normal((n, r)).astype(np.float32)
    return X1, y1, X2, y2, A


@pytest.mark.parametrize("m,n,d,r", [(300, 257, 18, 40), (1000, 1000, 18, 32), (513, 700, 54, 100),
                                     (129, 64, 32, 256), (77, 33, 5, 7), (256, 2048, 18, 512)])
def test_kprod_fp32(ctx, m, n, d, r):
    import tensorsvm_amd as T
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, m + n)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_FP32)
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.05)
    scale = np.abs(ref).max()
    # FP32 products accumulated over n terms: relative to the output scale
    assert np.abs(W - ref).max() <= 2e-5 * scale * max(1.0, np.sqrt(n / 256))


@pytest.mark.parametrize("m,n,d,r", [(300, 257, 18, 40), (1000, 1000, 18, 32), (513, 700, 54, 100),
                                     (129, 64, 32, 256), (77, 33, 5, 7), (256, 2048, 18, 512)])
def test_kprod_bf16(ctx, m, n, d, r):
    import tensorsvm_amd as T
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, m + n + 1)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_BF16)
    # oracle on the bf16-rounded right-hand side: the remaining error is the bf16
    # rounding of the kernel values (relative 2^-9, unbiased) and the split distance
    import torch
    Ab = torch.from_numpy(A).to(torch.bfloat16).to(torch.float32).numpy()
    ref = O.kernel_product(X1, y1, X2, y2, Ab, 0.05)
    # E = bf16(K) = K (1 + delta), |delta| <= 2^-9 (round to nearest), so the error of an
    # entry is bounded by 2^-9 sum_j |K_ij A_jc| (plus FP32 accumulation and the ~2^-16
    # relative error of the split distance), and on average far below that bound.
    Kabs = O.kernel_product(X1, np.ones_like(y1), X2, np.ones_like(y2), np.abs(Ab), 0.05)
    err = np.abs(W - ref)
    bound = 2.0 ** -9 * Kabs
    assert (err <= 1.05 * bound + 1e-5 * np.abs(ref).max()).all()
    assert np.median(err / (bound + 1e-30)) < 0.3


def test_kprod_gaussian_mixture_gamma(ctx):
    """Config-shaped data (d=18, gamma=0.05) and a larger gamma; FP32 form vs FP64."""
    import tensorsvm_amd as T
    X, y = gaussian_mixture(3000, 18, 1000)
    rng = np.random.default_rng(5)
    A = rng.standard_normal((3000, 64)).astype(np.float32)
    for g in (0.05, 1.0):
        W = ctx.kernel_product(X, y, X, y, A, g, T.PREC_FP32)
        ref = O.kernel_product(X, y, X, y, A, g)
        assert np.abs(W - ref).max() <= 3e-5 * np.abs(ref).max()


@pytest.mark.parametrize("m,n,d,r", [(300, 257, 18, 40), (1000, 1000, 18, 32), (513, 700, 54, 100),
                                     (129, 64, 32, 256), (77, 33, 5, 7), (256, 2048, 18, 512)])
def test_kprod_fp16x2(ctx, m, n, d, r):
    """Split form: kernel values as fp16 hi + lo (about 22 bits), the right-hand side
    rounded to fp16 under one power-of-two scale (relative 2^-11), distance as in BF16."""
    import tensorsvm_amd as T
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, m + n + 2)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_FP16X2)
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.05)
    Kabs = O.kernel_product(X1, np.ones_like(y1), X2, np.ones_like(y2), np.abs(A), 0.05)
    err = np.abs(W - ref)
    bound = (2.0 ** -11 + 2.0 ** -13) * Kabs
    assert (err <= 1.05 * bound + 1e-6 * np.abs(ref).max()).all()
    # and well inside the plain BF16 form's error on the same inputs
    Wb = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_BF16)
    assert np.median(err) < 0.35 * np.median(np.abs(Wb - ref))


@pytest.mark.parametrize("m,n,d,r", [(300, 257, 18, 40), (1000, 1000, 18, 100), (513, 700, 54, 100),
                                     (1100, 3001, 18, 300), (129, 64, 32, 256), (77, 33, 5, 7),
                                     (2048, 2048, 18, 512), (40, 5000, 18, 128),
                                     (1100, 3001, 41, 300), (700, 900, 43, 200),
                                     (300, 700, 64, 200), (300, 700, 84, 130)])
def test_kprod_fp16(ctx, m, n, d, r):
    """Plain fp16 form (the factor products of the BF16 training form): kernel values and
    the scaled right-hand side each rounded once to fp16 (relative 2^-11), distance as in
    BF16; ragged row, column and right-hand-side counts, several column tiles."""
    import tensorsvm_amd as T
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, m + n + 3)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_FP16)
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.05)
    Kabs = O.kernel_product(X1, np.ones_like(y1), X2, np.ones_like(y2), np.abs(A), 0.05)
    err = np.abs(W - ref)
    bound = (2.0 ** -10 + 2.0 ** -13) * Kabs
    assert (err <= 1.05 * bound + 1e-6 * np.abs(ref).max()).all()
    assert np.median(err / (bound + 1e-30)) < 0.3


def test_kprod_fp16x2_exact_rhs(ctx):
    """With A = I (exact in fp16) every output is one labelled kernel value, so only the
    split of E and the distance remain: the per-entry error must sit well below one fp16
    rounding of E (median 2^-12 relative), which a lost low term would leave."""
    import tensorsvm_amd as T
    m, n, d = 700, 96, 18
    X1, y1, X2, y2, _ = _case(m, n, d, n, 0.05, 99)
    A = np.eye(n, dtype=np.float32)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_FP16X2)
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.05)
    rel = np.abs(W - ref) / np.abs(ref)
    assert np.median(rel) < 2.0 ** -14
    assert rel.max() < 2.0 ** -12


@pytest.mark.parametrize("prec,d", [("bf16", 18), ("fp16", 18), ("fp16", 2), ("fp16", 9),
                                    ("fp16", 14), ("bf16", 25), ("fp16", 30), ("fp16", 36),
                                    ("fp16", 41)])
def test_kprod_wide_tiles_match_narrow(ctx, prec, d):
    """Right-hand sides wider than 128 columns run the 256-column-tile kernel; its outputs
    are bit-identical to the 128-column kernel's on each half (same distance tiles, same
    accumulation order).  For fp16 both halves get the same power-of-two scale by carrying
    the same largest |A| entry.  The feature counts cover every compiled operand width of
    the wide kernel (KD = 16 .. 128: one to eight distance K steps)."""
    import tensorsvm_amd as T
    p = T.PREC_BF16 if prec == "bf16" else T.PREC_FP16
    m, n, r = 1000, 3000, 256
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, 123)
    A[0, 0], A[0, 128] = 7.5, 7.5
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, p)
    W0 = ctx.kernel_product(X1, y1, X2, y2, np.ascontiguousarray(A[:, :128]), 0.05, p)
    W1 = ctx.kernel_product(X1, y1, X2, y2, np.ascontiguousarray(A[:, 128:]), 0.05, p)
    assert np.array_equal(W[:, :128], W0)
    assert np.array_equal(W[:, 128:], W1)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_kprod_wide_xcd_schedule_matches_narrow(ctx, prec):
    """With at least two rounds of 256-row x 256-column tiles (here 547 row blocks x 2
    column tiles) the wide kernel runs as one cooperative launch whose blocks take units of
    tiles per XCD from a device-wide counter; every tile is computed as in the one launch per
    round, so each 128-column quarter is bit-identical to the 128-column kernel's."""
    import tensorsvm_amd as T
    p = T.PREC_BF16 if prec == "bf16" else T.PREC_FP16
    m, n, d, r = 140_000, 2000, 18, 512
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, 321)
    for j in range(4):
        A[0, 128 * j] = 7.5
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, p)
    for j in range(4):
        Wj = ctx.kernel_product(X1, y1, X2, y2, np.ascontiguousarray(A[:, 128 * j:128 * (j + 1)]),
                                0.05, p)
        assert np.array_equal(W[:, 128 * j:128 * (j + 1)], Wj), j


@pytest.mark.parametrize("m,n,d,r", [(300, 257, 18, 40), (1000, 1000, 18, 512), (513, 700, 54, 100),
                                     (129, 64, 32, 256), (77, 33, 5, 7), (200, 3001, 41, 130)])
def test_kprod_fp64(ctx, m, n, d, r):
    """FP64 form: distance, exp and product in double; only the FP32 output rounds."""
    import tensorsvm_amd as T
    X1, y1, X2, y2, A = _case(m, n, d, r, 0.05, m + n + 4)
    W = ctx.kernel_product(X1, y1, X2, y2, A, 0.05, T.PREC_FP64)
    ref = O.kernel_product(X1, y1, X2, y2, A, 0.05)
    err = np.abs(W - ref)
    assert (err <= 2.0 ** -24 * np.abs(ref) + 1e-12 * np.abs(ref).max()).all(), err.max()
