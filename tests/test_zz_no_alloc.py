"""tsvm_kernel_product_noalloc_dev (the allocation-free product) through the C-ABI against
tsvm_kernel_product_dev on the same device inputs: identical outputs (the same kernels) in
every arithmetic form, on two calls in a row with the second smaller (the shared scratch
keeps the first call's contents), and a clear argument error beyond the scratch size."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _inputs(m, n, d, r, seed):
    rng = np.random.default_rng(seed)
    X1 = (rng.standard_normal((m, d)) * 0.7).astype(np.float32)
    X2 = (rng.standard_normal((n, d)) * 0.7).astype(np.float32)
    y1 = np.where(rng.random(m) < 0.5, 1.0, -1.0).astype(np.float32)
    y2 = np.where(rng.random(n) < 0.5, 1.0, -1.0).astype(np.float32)
    A = rng.standard_normal((n, r)).astype(np.float32)
    return [_dev(a) for a in (X1, y1, X2, y2, A)]


@pytest.mark.parametrize("prec", [0, 1, 2, 3, 4])
def test_noalloc_matches_workspace_product(ctx, prec):
    import torch
    # the second call is smaller in every dimension; the FP32 form's 9000 columns take
    # two 8192-column segments (partial sums in the scratch) in the first call
    for (m, n, d, r, seed) in [(1500, 9000, 18, 96, 1), (300, 700, 7, 33, 2)]:
        X1, y1, X2, y2, A = _inputs(m, n, d, r, seed + 10 * prec)
        W_ws = torch.full((m, r), float("nan"), dtype=torch.float32, device="cuda")
        W_na = torch.full((m, r), float("nan"), dtype=torch.float32, device="cuda")
        ctx.kernel_product_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, prec, W_ws)
        ctx.kernel_product_noalloc_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, prec, W_na)
        ctx.sync()
        a, b = W_ws.cpu().numpy(), W_na.cpu().numpy()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (m, n, prec, np.abs(a - b).max())


@pytest.mark.parametrize("prec", [1, 3])
def test_noalloc_matches_workspace_product_wide(ctx, prec):
    """Right-hand sides wider than 128 columns (the 256-column-tile kernel of the bf16 and
    fp16 forms) through the scratch array: two calls, the second smaller."""
    import torch
    for (m, n, d, r, seed) in [(800, 6000, 18, 256, 3), (300, 2000, 18, 160, 4)]:
        X1, y1, X2, y2, A = _inputs(m, n, d, r, seed + 10 * prec)
        W_ws = torch.full((m, r), float("nan"), dtype=torch.float32, device="cuda")
        W_na = torch.full((m, r), float("nan"), dtype=torch.float32, device="cuda")
        ctx.kernel_product_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, prec, W_ws)
        ctx.kernel_product_noalloc_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, prec, W_na)
        ctx.sync()
        a, b = W_ws.cpu().numpy(), W_na.cpu().numpy()
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), (m, n, prec, np.abs(a - b).max())


def test_noalloc_rejects_oversize(ctx):
    import tensorsvm_amd as T
    # 2 n r bytes of fp16 right-hand side alone exceed the 64 MiB scratch: refused before
    # any work is queued (the buffers are never touched)
    m, n, d, r = 256, 70000, 18, 512
    X1, y1, X2, y2, A = _inputs(m, n, d, r, 3)
    import torch
    W = torch.zeros((m, r), dtype=torch.float32, device="cuda")
    with pytest.raises(T.capi.TsvmError, match="status 1.*scratch"):
        ctx.kernel_product_noalloc_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, 3, W)
    # and the entry point still works afterwards
    m, n, r = 64, 100, 8
    X1, y1, X2, y2, A = _inputs(m, n, d, r, 4)
    Wa = torch.zeros((m, r), dtype=torch.float32, device="cuda")
    Wb = torch.zeros((m, r), dtype=torch.float32, device="cuda")
    ctx.kernel_product_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, 0, Wa)
    ctx.kernel_product_noalloc_dev(X1, y1, m, X2, y2, n, d, A, r, 0.05, 0, Wb)
    ctx.sync()
    assert torch.equal(Wa, Wb)
