"""Low-rank factor construction (HIP, through the C-ABI) against the oracle's LRA."""
import numpy as np
import pytest
from scipy.spatial.distance import cdist

import oracle_lib as O
from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu


def _ky(X, y, gamma):
    X = X.astype(np.float64)
    return y[:, None] * np.exp(-gamma * cdist(X, X, "sqeuclidean")) * y[None, :]


@pytest.mark.parametrize("n,d,k,gamma", [(2000, 18, 32, 0.05), (1500, 54, 64, 1.0), (3000, 18, 100, 0.05)])
def test_lra_fp32_matches_oracle(ctx, n, d, k, gamma):
    import tensorsvm_amd as T
    X, y = gaussian_mixture(n, d, 1000 + n)
    om = O.gaussrand(n * k).reshape(n, k)
    U, info = ctx.lra(X, y, k, gamma, T.PREC_FP32, omega=om)
    Uo, w, realk = O.lra(X, y, k, gamma, om)
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    Go = Uo @ Uo.T
    rel = np.linalg.norm(G - Go) / np.linalg.norm(Go)
    assert rel < 1e-4, rel
    # the approximation error against the exact kernel agrees too
    K = _ky(X, y, gamma)
    e_gpu = np.linalg.norm(K - G) / np.linalg.norm(K)
    e_orc = np.linalg.norm(K - Go) / np.linalg.norm(K)
    assert abs(e_gpu - e_orc) <= 1e-3 * max(e_orc, 1e-3)
    assert info.realk == realk or abs(info.realk - realk) <= 1
    assert np.isclose(info.eig_max, w[-1], rtol=1e-4)


def test_lra_full_rank_reproduces_kernel(ctx):
    """k = n with distinct points: range(K_Y Omega) is everything, so U U^T = K_Y."""
    import tensorsvm_amd as T
    n, d, gamma = 256, 8, 0.3
    X, y = gaussian_mixture(n, d, 7)
    om = O.gaussrand(n * n).reshape(n, n)
    U, info = ctx.lra(X, y, n, gamma, T.PREC_FP32, omega=om)
    K = _ky(X, y, gamma)
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    assert np.linalg.norm(K - G) / np.linalg.norm(K) < 1e-4


@pytest.mark.parametrize("form,same_tol", [("split", 1e-3), ("bf16", 1e-2)])
def test_lra_low_precision_close_to_fp64(ctx, form, same_tol):
    """Low-precision forms: the same projected kernel as the FP64 method.  The basis stays
    FP32 and the second product runs in fp16; the sketch runs in the two-term product
    (split: span(W) matches to ~fp16 rounding of Omega) or in the fp16 product (bf16: each
    kernel value of the sketch rounded once to fp16, 2^-11)."""
    import tensorsvm_amd as T
    prec = T.PREC_FP16X2 if form == "split" else T.PREC_BF16
    n, d, k, gamma = 4000, 18, 64, 0.05
    X, y = gaussian_mixture(n, d, 99)
    om = O.gaussrand(n * k).reshape(n, k)
    U, info = ctx.lra(X, y, k, gamma, prec, omega=om)
    Uo, w, realk = O.lra(X, y, k, gamma, om)
    K = _ky(X, y, gamma)
    G = U.astype(np.float64) @ U.T.astype(np.float64)
    Go = Uo @ Uo.T
    e_gpu = np.linalg.norm(K - G) / np.linalg.norm(K)
    e_orc = np.linalg.norm(K - Go) / np.linalg.norm(K)
    # the low-rank approximation is as good as the FP64 one, to within 2% of its error
    assert e_gpu <= 1.02 * e_orc + 2e-3, (e_gpu, e_orc)
    # and it is the same approximation, not merely an equally good one
    e_same = np.linalg.norm(G - Go) / np.linalg.norm(Go)
    print(f"{form} form: |K-UU'|/|K| {e_gpu:.3e} (fp64 {e_orc:.3e}), |UU' - UoUo'|/|UoUo'| {e_same:.3e}")
    assert e_same < same_tol * e_orc, (e_same, e_orc)


@pytest.mark.parametrize("n,d,k,gamma", [(1500, 18, 64, 0.05), (700, 54, 40, 1.0)])
def test_fnorm_res_matches_oracle(ctx, n, d, k, gamma):
    """-analysis on the GPU: |K_Y|_F and |K_Y - U U^T|_F against the FP64 oracle on the same U."""
    import torch
    import tensorsvm_amd as T
    X, y = gaussian_mixture(n, d, 13 + n)
    om = O.gaussrand(n * k).reshape(n, k)
    U, info = ctx.lra(X, y, k, gamma, T.PREC_FP32, omega=om)
    fo, ro = O.fnorm_res(X, y, gamma, U.astype(np.float64))
    dX, dy, dU = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (X, y, U))
    f, r = ctx.fnorm_res_dev(dX, dy, n, d, gamma, dU, k)
    assert abs(f - fo) <= 1e-6 * fo
    assert abs(r - ro) <= 1e-5 * ro + 1e-6 * fo
    # through the trainer's -analysis flag
    al, _, ti = ctx.train(X, y, k=k, C=4.0, gamma=gamma, precision=T.PREC_FP32, omega=om, analysis=1)
    assert abs(ti.fnorm - fo) <= 1e-6 * fo and abs(ti.fnorm_res - ro) <= 1e-5 * ro + 1e-6 * fo

