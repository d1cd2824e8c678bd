"""GPU interior-point solver (through the C-ABI) against the oracle's mpc and sklearn."""
import numpy as np
import pytest

import oracle_lib as O
from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu


def _problem(n, k, seed, sep=0.05):
    rng = np.random.default_rng(seed)
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    U = rng.normal(size=(n, k)) * 0.6 + sep * y[:, None] * rng.normal(size=(1, k))
    Z = (y[:, None] * U).astype(np.float32)
    return Z, y.astype(np.float32)


@pytest.mark.parametrize("name", ["a", "b"])
def test_ipm_fp32_vs_sklearn_fixture(ctx, name):
    import tensorsvm_amd as T
    f = np.load(f"tests/golden/ipm_svc_{name}.npz")
    C = float(f["C"])
    Z = f["Z"].astype(np.float32)
    al, xi, s, info = ctx.ipm(Z, f["y"], C, precision=T.PREC_FP32)
    assert info.reason == 0, info.as_dict()
    assert np.abs(al - f["alpha"]).max() <= 1e-4 * C
    assert abs(info.y - (-float(f["b"]))) <= 1e-3 * (1 + abs(float(f["b"])))
    # and the oracle on the identical FP32 factor agrees more tightly
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), f["y"], C)
    assert np.abs(al - Xo).max() <= 1e-6 * C
    assert info.iters == io.iters


# k < 256 runs the narrow pass kernels; 256 (one float4 group per lane), 300 and 600
# (a partly masked last group), 1000 (four groups, separate residual pass) the wide ones
@pytest.mark.parametrize("n,k,C", [(5000, 32, 32.0), (3001, 50, 1.0), (8000, 100, 8.0),
                                   (4000, 160, 8.0), (6000, 200, 32.0), (3000, 72, 4.0),
                                   (4000, 256, 8.0), (4000, 300, 32.0), (3000, 600, 8.0),
                                   (1200, 1000, 8.0)])
@pytest.mark.parametrize("refine", [1, -1])
def test_ipm_fp32_matches_oracle(ctx, n, k, C, refine):
    import tensorsvm_amd as T
    Z, y = _problem(n, k, n + k)
    al, xi, s, info = ctx.ipm(Z, y, C, precision=T.PREC_FP32, refine=refine)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y.astype(np.float64), C)
    assert info.reason == io.reason
    assert np.abs(al - Xo).max() <= 1e-5 * C
    assert abs(info.pobj - io.pobj) <= 1e-7 * (1 + abs(io.pobj))
    # same support-vector sets apart from multipliers within 1e-6 of 0 or C
    sv_g = al > 1e-6
    sv_o = Xo > 1e-6
    amb = (np.abs(Xo) < 1e-5) | (np.abs(Xo - C) < 1e-5)
    assert np.all((sv_g == sv_o) | amb)


# refine = 2: two sweeps for each of the affine and the corrector's steps; -1: one sweep
# for the combined step only (the affine step as solved)
@pytest.mark.parametrize("refine", [2, -1])
@pytest.mark.parametrize("n,k,C", [(5000, 32, 32.0), (8000, 100, 8.0), (6000, 200, 32.0),
                                   (4000, 300, 32.0), (3000, 600, 8.0)])
def test_ipm_bf16_refined_matches_oracle(ctx, n, k, C, refine):
    import tensorsvm_amd as T
    Z, y = _problem(n, k, 7 * n + k)
    al, xi, s, info = ctx.ipm(Z, y, C, precision=T.PREC_BF16, refine=refine,
                              stop=T.capi.STOP_GAP, tol=1e-7)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y.astype(np.float64), C, stop="gap", tol=1e-7)
    assert info.reason == 0, info.as_dict()
    # the hybrid capacitance put the bulk rows on BF16 in the later iterations
    assert info.syrk_bf16_iters > 0 and info.bf16_row_frac > 0, info.as_dict()
    assert abs(info.pobj - io.pobj) <= 1e-6 * (1 + abs(io.pobj))
    assert np.abs(al - Xo).max() <= 1e-4 * C


def test_ipm_linear_init(ctx):
    """kernel_type 0 (linear SVM): Z = diag(y) X, the reference's linear initial point."""
    import tensorsvm_amd as T
    rng = np.random.default_rng(3)
    n, d = 2000, 10
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    X = rng.normal(size=(n, d)) + 0.5 * y[:, None]
    Z = (y[:, None] * X).astype(np.float32)
    al, xi, s, info = ctx.ipm(Z, y, 1.0, kernel_type=0, precision=T.PREC_FP32)
    Xo, Xio, So, io = O.mpc(Z.astype(np.float64), y, 1.0, kernel_type=0)
    assert info.reason == io.reason
    assert np.abs(al - Xo).max() <= 1e-5


@pytest.mark.parametrize("n,k,C", [(5000, 32, 32.0), (8000, 100, 8.0)])
def test_ipm_feasible_init_same_optimum(ctx, n, k, C):
    """The zero-dual-residual start reaches the same optimum as the reference start."""
    import tensorsvm_amd as T
    Z, y = _problem(n, k, 11 * n + k)
    al0, _, _, i0 = ctx.ipm(Z, y, C, stop=T.capi.STOP_GAP, tol=1e-7)
    al1, _, _, i1 = ctx.ipm(Z, y, C, stop=T.capi.STOP_GAP, tol=1e-7, init=1)
    assert i0.reason == 0 and i1.reason == 0
    assert abs(i0.pobj - i1.pobj) <= 1e-6 * (1 + abs(i0.pobj))
    assert np.abs(al0 - al1).max() <= 1e-3 * C


def test_reference_typed_entry_points(ctx):
    """tsvm_lra_f64 / tsvm_mpc take the reference's types (double U with a leading
    dimension, LRA's return value; double Z and labels for mpc, tensorsvm-train.cu:100,
    111): the same factor widened exactly, the largest eigenvalue as l1, and the same
    solve as the FP32-factor entry points, bit for bit."""
    import tensorsvm_amd as T
    X, y = gaussian_mixture(3000, 6, 41)
    n, k, gamma, C = 3000, 40, 0.1, 4.0
    om = O.gaussrand(n * k)
    U32, i32 = ctx.lra(X, y, k, gamma, T.PREC_FP32, omega=om)
    U64, l1, i64 = ctx.lra_f64(X, y, k, gamma, T.PREC_FP32, omega=om, ldu=k + 3)
    assert np.array_equal(U64[:, :k], U32.astype(np.float64))
    assert not U64[:, k:].any()  # beyond k columns of each row: untouched
    assert l1 == i32.eig_max == i64.eig_max
    al, xi, _, ii = ctx.ipm(U32, y, C)
    X64, Xi64, mi = ctx.mpc(U64, y.astype(np.float64), C, k=k)
    assert mi.iters == ii.iters and mi.reason == ii.reason
    assert np.array_equal(X64, al) and np.array_equal(Xi64, xi)
