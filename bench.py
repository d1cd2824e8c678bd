"""Benchmark of the TensorSVM training hot path on one MI355X.

Workload (the precision-form configuration the metric is quoted on): 2,000,000 rows of
18-feature Gaussian-mixture data (class means +-0.5 on every dimension, unit variance,
seeded), 80/20 split -> 1,600,000 training rows and 400,000 held-out rows; RBF gamma=0.05,
C=32, rank k=512; training stops at a relative duality gap of 1e-5 (with the primal and
dual feasibility tests of the gap rule).

A step is one full training run (factor construction + interior-point solve + bias) in
the BF16 form, with X, y and Omega already resident in HBM (Omega, the reference's
gaussrand() stream, is generated and uploaded outside the timed region).  The plain FP32
form is run once after the timed steps for the comparison the metric names, and both
models are scored on the held-out rows.

Reported beside the timing:
  roofline      the dominant kernel, the factor's sketch product K_Y Omega (split-fp16
                fused kernel product): algorithmic flops 2 m n (d + k) per launch / its
                launch time (HIP events on the library's stream), against the dense
                BF16/FP16 MFMA peak; traffic = HBM bytes per launch from rocprofv3
                FETCH_SIZE + WRITE_SIZE (calibrated on a copy of known size)
  zdz           Z^T D^-1 Z: the BF16 MFMA part of the hybrid capacitance (n k (k+1) flops
                per call / its average time, HIP events) against the BF16 peak
  grad          the differentiation form (200,000 + 50,000 rows, d = 32, k = 256, FP32):
                gradient of the trained dual and of the held-out smoothed hinge in
                (log gamma, 32 log-scales), backward flops / time against the FP32 peak,
                with a central-difference check of the log-gamma component (h = 1e-3)
  cpu_baseline  the same algorithm on the host cores through a CPU BLAS (oracle/cpu_blas.py:
                the reference's FP32 tile pipeline for the kernel products, its FP64
                interior-point solver), timed on bounded samples and extrapolated: one
                product on a 50,000 x 50,000 slice scaled by (n / 50,000)^2, two of them;
                solver iterations on 200,000 rows of a factor-shaped matrix scaled by
                n / 200,000 and by the GPU run's iteration count
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "IPM train s to gap 1e-5 (n=2M,k=512): BF16+refine vs FP32; Z^T*D*Z MFMA %"
PEAK_BF16 = 2500.0   # TFLOP/s, dense
PEAK_FP32 = 157.3
PEAK_FP64 = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# HBM bytes per product of the factor's product kernel at the default workload (mean of its
# two products, one dispatch each): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# (KiB per dispatch) scaled by the factor that the same counters give for the bench's
# 3.28 GB Omega copy (x2.0 and x1.0)
KPROD_TRAFFIC = {"fetch_bytes": 401_741_402_112, "write_bytes": 3_304_685_568,
                 "source": "profiles/kprod_factor_c2_pmc.md"}


def grad_form(ctx, torch, n=200_000, nt=50_000, d=32, k=256, C=1.0, seed=1003):
    """Differentiation form: train (FP32), then the gradient of the trained dual and of
    the held-out smoothed hinge; returns the measurement object."""
    import numpy as np
    import tensorsvm_amd as T
    from datagen import scaled_mixture
    X, y, g0 = scaled_mixture(n + nt, d, seed)
    Xt, yt = np.ascontiguousarray(X[n:]), np.ascontiguousarray(y[n:])
    X, y = np.ascontiguousarray(X[:n]), np.ascontiguousarray(y[:n])
    dev = torch.device("cuda:0")
    dX, dy = torch.from_numpy(X).to(dev), torch.from_numpy(y).to(dev)
    dXt, dyt = torch.from_numpy(Xt).to(dev), torch.from_numpy(yt).to(dev)
    om = torch.from_numpy(T.gaussrand_fill(n * k)).to(dev)
    dom = om.clone()
    dal = torch.empty(n, dtype=torch.float64, device=dev)
    prm = T.capi.train_params(k=k, C=C, gamma=g0, kernel_type=2, precision=T.PREC_FP32,
                              stop=T.capi.STOP_GAP, tol=1e-5, maxit=300)
    t0 = time.perf_counter()
    info = ctx.train_dev(dX, dy, n, d, prm, dom, dal)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    lg, s0 = float(np.log(g0)), np.zeros(d)
    D, L, gD, gL, gi = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg, s0, om, dXt, dyt, nt)
    h = 1e-3
    Dp, Lp, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg + h, s0, om, dXt, dyt, nt, False)
    Dm, Lm, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg - h, s0, om, dXt, dyt, nt, False)
    fdD, fdL = (Dp - Dm) / (2 * h), (Lp - Lm) / (2 * h)
    h2 = 1e-2  # a wider step: FP32 noise of the objectives / h shrinks tenfold
    Dp2, Lp2, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg + h2, s0, om, dXt, dyt, nt, False)
    Dm2, Lm2, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg - h2, s0, om, dXt, dyt, nt, False)
    fdD2, fdL2 = (Dp2 - Dm2) / (2 * h2), (Lp2 - Lm2) / (2 * h2)
    # Richardson: central differences at 2h and h, (4 FD(h) - FD(2h)) / 3 cancels the h^2
    # term of the truncation error
    Dp3, Lp3, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg + 2 * h, s0, om, dXt, dyt, nt, False)
    Dm3, Lm3, _, _, _ = ctx.grad_dev(dX, dy, n, d, dal, info.b, k, lg - 2 * h, s0, om, dXt, dyt, nt, False)
    fdD3, fdL3 = (Dp3 - Dm3) / (4 * h), (Lp3 - Lm3) / (4 * h)
    rD, rL = (4 * fdD - fdD3) / 3, (4 * fdL - fdL3) / 3
    tf = gi.backward_tflops
    log(f"grad: train {t_train:.2f}s, gradient {gi.t_total:.2f}s (backward {gi.t_backward:.3f}s, "
        f"{tf:.1f} TF/s); dD/dlog(gamma) {gD[0]:.6e} vs FD {fdD:.6e}")
    return {"workload": f"scaled-variance mixture n={n} train + {nt} held-out, d={d}, k={k}, FP32",
            "t_train_s": round(t_train, 3), "t_grad_s": round(gi.t_total, 3),
            "t_backward_s": round(gi.t_backward, 4), "backward_tflops": round(tf, 2),
            "t_backward_heldout_s": round(gi.t_backward_heldout, 4),
            "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": PEAK_FP32,
                         "unit": "TFLOP/s", "frac": round(tf / PEAK_FP32, 4)},
            "dual": D, "heldout_loss": L,
            "dD_dloggamma": gD[0], "dD_dloggamma_fd": fdD,
            "dD_rel_err_vs_fd": abs(gD[0] - fdD) / abs(fdD),
            "dL_dloggamma": gL[0], "dL_dloggamma_fd": fdL,
            "dL_rel_err_vs_fd": abs(gL[0] - fdL) / max(abs(fdL), 1e-300),
            "fd_h1e-2": {"dD_rel_err": abs(gD[0] - fdD2) / abs(fdD2),
                         "dL_rel_err": abs(gL[0] - fdL2) / max(abs(fdL2), 1e-300)},
            "fd_richardson_h1e-3": {"dD_rel_err": abs(gD[0] - rD) / abs(rD),
                                    "dL_rel_err": abs(gL[0] - rL) / max(abs(rL), 1e-300)},
            "ipm": {"iters": info.ipm.iters, "reason": info.ipm.reason,
                    "gap_rel": info.ipm.gap_rel}, "nsv": info.nsv}


def cpu_baseline(X, y, gamma, k, ntr, iters, prod_rows=50_000, ipm_rows=200_000):
    """The reference's plain method on the host cores with a CPU BLAS (MKL through torch's
    CPU kernels), timed on bounded samples of the workload and extrapolated."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_blas
    th = torch.get_num_threads()
    m = min(prod_rows, ntr)
    g = torch.Generator().manual_seed(3)
    Xs, ys = torch.from_numpy(X[:m]), torch.from_numpy(y[:m])
    A = torch.randn(m, k, generator=g)
    t0 = time.perf_counter()
    cpu_blas.kernel_product_tiles(Xs, ys, Xs, ys, A, gamma)
    tp = time.perf_counter() - t0
    prod_est = tp * (ntr / m) ** 2 * 2.0
    del A
    ni = min(ipm_rows, ntr)
    Z = torch.randn(ni, k, generator=g, dtype=torch.float64) * 0.05
    a = torch.where(torch.rand(ni, generator=g) < 0.5, 1.0, -1.0).double()
    t0 = time.perf_counter()
    cpu_blas.mpc(Z, a, 32.0, iter_limit=0)
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    cpu_blas.mpc(Z, a, 32.0, iter_limit=2)
    t_it = (time.perf_counter() - t0 - t_init) / 2.0
    ipm_est = t_it * (ntr / ni) * iters
    log(f"cpu baseline: product sample {tp:.1f}s -> {prod_est:.0f}s, ipm {t_it:.2f}s/it -> {ipm_est:.0f}s")
    return {"value": round(prod_est + ipm_est, 1), "unit": "s", "cores": th, "kind": "port",
            "extrapolated": True, "blas": "MKL through torch CPU kernels (sgemm, dgemm, potrf)",
            "factor_products_s": round(prod_est, 1), "ipm_s": round(ipm_est, 1),
            "sample": f"the reference's FP32 tile pipeline (8192^2 tiles: XIJ sgemm, exp, K A "
                      f"sgemm) for one product over a {m} x {m} slice, k={k}: {tp:.2f}s, "
                      f"x (n/{m})^2 x 2 products; its FP64 IPM (dsyrk-as-GEMM capacitance, "
                      f"potrf, SMW gemv) {t_it:.3f}s per iteration on {ni} x {k}, x n/{ni} x "
                      f"{iters} iterations (the GPU run's count); {th} threads"}


def dense_pass_figures(ipm, n, k):
    """Per-iteration time and rates of the capacitance's dense pass (BF16 MFMA)."""
    t, it = ipm.get("t_syrk_dense", 0.0), ipm.get("syrk_bf16_iters", 0)
    if not t or not it:
        return {"dense_pass_ms_per_iter": None}
    per = t / it
    syrk = n * k * (k + 1) / per / 1e12
    return {"dense_pass_ms_per_iter": round(1e3 * per, 3),
            "dense_pass_mfma_tflops": round(ipm["syrk_dense_mfma_tflops"], 1),
            "dense_pass_mfma_frac_of_bf16_peak": round(ipm["syrk_dense_mfma_tflops"] / PEAK_BF16, 4),
            "dense_pass_syrk_tflops": round(syrk, 1),
            "dense_pass_syrk_frac_of_bf16_peak": round(syrk / PEAK_BF16, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=2_000_000, help="rows generated (80%% train)")
    ap.add_argument("--k", type=int, default=512)
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-grad", action="store_true", help="skip the differentiation form")
    ap.add_argument("--data", choices=("gaussian", "covtype"), default="gaussian",
                    help="gaussian: the mixture of the default workload (d=18); covtype: the "
                         "covtype-like set (d=54, 10 uniform + 44 binary columns, RBF labels)")
    ap.add_argument("--gamma", type=float, default=None, help="RBF gamma (0.05 gaussian, 1.0 covtype)")
    ap.add_argument("--analysis", action="store_true", help="|K_Y - U U^T|_F / |K_Y|_F after the factor")
    ap.add_argument("--grad-timed", action="store_true",
                    help="gradient (log gamma and per-feature scales) of the timed model at its own "
                         "size, forward products in the timed form's precision")
    ap.add_argument("--progress", action="store_true",
                    help="progress lines on stderr from the library (row slabs of each product)")
    ap.add_argument("--fp64-ref", action="store_true",
                    help="train once more with FP64 factor products (outside the timed steps) and "
                         "state the timed form's and the FP32 form's error against it")
    ap.add_argument("--precision", choices=("bf16", "fp32", "fp8"), default="bf16",
                    help="form of the timed steps (the comparison run, if any, is FP32); fp8: the "
                         "BF16 form with the FP8 e4m3 pass of the capacitance")
    ap.add_argument("--refine", type=int, default=-1,
                    help="refinement sweeps of the Newton steps (m > 0: m for the affine step and m "
                         "for the corrector's, each solved first; -m: m for the combined step, "
                         "started from the affine one, the first adding the corrector's step)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the trainer "
              "runs on one GPU and does not shard its work across GPUs", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import tensorsvm_amd as T
    from datagen import covtype_like, gaussian_mixture

    C, k = 32.0, args.k
    t = time.time()
    if args.data == "gaussian":
        d, gamma, seed = 18, 0.05, 1002
        Xall, yall = gaussian_mixture(args.n, d, seed)
    else:
        d, gamma, seed = 54, 1.0, 1001
        Xall, yall = covtype_like(args.n, seed)
    if args.gamma is not None:
        gamma = args.gamma
    rng = np.random.default_rng(seed + 1)
    perm = rng.permutation(args.n)
    ntr = int(0.8 * args.n)
    tr, te = perm[:ntr], perm[ntr:]
    X, y = np.ascontiguousarray(Xall[tr]), np.ascontiguousarray(yall[tr])
    Xte, yte = np.ascontiguousarray(Xall[te]), np.ascontiguousarray(yall[te])
    log(f"data: {ntr} train / {args.n - ntr} held-out, d={d} ({time.time() - t:.1f}s)")
    t = time.time()
    om = T.gaussrand_fill(ntr * k)
    t_omega = time.time() - t
    log(f"omega: {ntr}x{k} gaussrand stream in {t_omega:.1f}s")

    ctx = T.Context(0)
    if args.progress:
        ctx.set_progress(True)
    dev = torch.device("cuda:0")
    dX = torch.from_numpy(X).to(dev)
    dy = torch.from_numpy(y).to(dev)
    dOm0 = torch.from_numpy(om).to(dev)
    del om
    dOm = torch.empty_like(dOm0)
    dal = torch.empty(ntr, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def run(precision):
        dOm.copy_(dOm0)  # the factor construction consumes Omega
        torch.cuda.synchronize()
        prm = T.capi.train_params(k=k, C=C, gamma=gamma, kernel_type=2, precision=precision,
                                  stop=T.capi.STOP_GAP, tol=1e-5, maxit=300, refine=args.refine,
                                  progress=int(args.progress), analysis=int(args.analysis))
        t0 = time.perf_counter()
        info = ctx.train_dev(dX, dy, ntr, d, prm, dOm, dal)
        torch.cuda.synchronize()
        # the optional -analysis pass is a diagnostic, not part of the training time
        return time.perf_counter() - t0 - info.t_analysis, info

    def accuracy(info):
        alpha = dal.cpu().numpy()
        f = ctx.decision_function(X, y, alpha, gamma, info.b, Xte)
        return float(np.mean(np.sign(f) == yte)), alpha

    prec = {"bf16": T.PREC_BF16, "fp32": T.PREC_FP32, "fp8": T.PREC_FP8}[args.precision]
    low = args.precision in ("bf16", "fp8")  # fp16/bf16 factor products
    for w in range(args.warmup):
        dt, _ = run(prec)
        log(f"warmup {w}: {dt:.2f}s")
    times, info = [], None
    for s in range(args.steps):
        dt, info = run(prec)
        times.append(dt)
        log(f"step {s}: {args.precision} train {dt:.2f}s (lra {info.lra.t_total:.2f}s, ipm {info.ipm.t_total:.2f}s, "
            f"{info.ipm.iters} it, reason {info.ipm.reason}, gap {info.ipm.gap_rel:.2e})")
    acc_bf16, alpha_bf16 = accuracy(info)
    bf16_info = info.as_dict()
    bf16_b = info.b
    value = float(np.mean(times))

    grad_timed = None
    if args.grad_timed:
        # the multipliers of the timed run are still in dal (the comparison run comes later)
        dXte, dyte = torch.from_numpy(Xte).to(dev), torch.from_numpy(yte).to(dev)
        gprec = T.PREC_BF16 if low else T.PREC_FP32
        torch.cuda.synchronize()
        tg0 = time.perf_counter()
        Dg, Lg, gDg, gLg, gi = ctx.grad_dev(dX, dy, ntr, d, dal, info.b, k, float(np.log(gamma)),
                                            np.zeros(d), dOm0, dXte, dyte, Xte.shape[0],
                                            precision=gprec)
        torch.cuda.synchronize()
        tg = time.perf_counter() - tg0
        grad_timed = {"workload": f"gradient of the timed model in (log gamma, {d} log-scales), "
                                  f"{ntr} train + {Xte.shape[0]} held-out rows",
                      "forward_products": "fp16" if low else "fp32",
                      "t_grad_s": round(tg, 3), "t_backward_s": round(gi.t_backward, 3),
                      "backward_tflops": round(gi.backward_tflops, 1),
                      "roofline": {"bound": "mfma", "achieved": round(gi.backward_tflops, 1),
                                   "peak": PEAK_FP32, "unit": "TFLOP/s",
                                   "frac": round(gi.backward_tflops / PEAK_FP32, 4)},
                      "dual": Dg, "heldout_loss": Lg, "dD_dloggamma": gDg[0], "dL_dloggamma": gLg[0]}
        log(f"grad of the timed model: {tg:.1f}s (backward {gi.t_backward:.1f}s, "
            f"{gi.backward_tflops:.1f} TF/s); dD/dlog(gamma) {gDg[0]:.6e}")
        del dXte, dyte

    fp32 = None
    if not args.no_fp32 and low:
        dt32, info32 = run(T.PREC_FP32)
        acc32, alpha32 = accuracy(info32)
        f32 = ctx.decision_function(X, y, alpha32, gamma, info32.b, Xte)
        f16 = ctx.decision_function(X, y, alpha_bf16, gamma, bf16_b, Xte)
        fp32 = {"train_s": dt32, "accuracy": acc32, "info": info32.as_dict(),
                "alpha_max_abs_diff_vs_bf16": float(np.abs(alpha32 - alpha_bf16).max()),
                "heldout_sign_agreement_vs_bf16": float(np.mean(np.sign(f32) == np.sign(f16)))}
        log(f"fp32 train {dt32:.2f}s (lra {info32.lra.t_total:.2f}s), accuracy {acc32:.4f}")

    fp64_ref = None
    if args.fp64_ref:
        dt64, info64 = run(T.PREC_FP64)
        acc64, alpha64 = accuracy(info64)
        f64 = ctx.decision_function(X, y, alpha64, gamma, info64.b, Xte)
        p64 = info64.ipm.pobj

        def err(alpha_x, b_x, ipm_x):
            fx = ctx.decision_function(X, y, alpha_x, gamma, b_x, Xte)
            return {"alpha_max_abs_diff_over_C": float(np.abs(alpha_x - alpha64).max() / C),
                    "alpha_mean_abs_diff_over_C": float(np.abs(alpha_x - alpha64).mean() / C),
                    "b_abs_diff": abs(b_x - info64.b),
                    "pobj_rel_err": abs(ipm_x["pobj"] - p64) / abs(p64),
                    "gap_rel": ipm_x["gap_rel"],
                    "heldout_sign_agreement": float(np.mean(np.sign(fx) == np.sign(f64)))}
        fp64_ref = {"form": "both factor products on the FP64 MFMA (factor stored in FP32), "
                            "FP64 capacitance, same stop rule",
                    "train_s": round(dt64, 2), "accuracy": acc64, "b": info64.b,
                    "pobj": p64, "gap_rel": info64.ipm.gap_rel, "iters": info64.ipm.iters,
                    "t_product1_s": round(info64.lra.t_product1, 2),
                    "timed_form_vs_fp64": err(alpha_bf16, bf16_b, bf16_info["ipm"]),
                    "fp32_vs_fp64": (err(alpha32, info32.b, fp32["info"]["ipm"]) if fp32 else None)}
        log(f"fp64 reference train {dt64:.1f}s, accuracy {acc64:.4f}")

    # roofline of the dominant kernel.  BF16 form: both products of the factor run the
    # same kernel (fp16 product), so its average product is the mean of the two; the
    # FP32 form likewise runs kprod_f32 twice.  The wide kernel (256-row x 256-column
    # tiles) runs a product with at least two rounds of tiles as one cooperative launch of
    # one block per CU on the XCD-round schedule, a smaller one as ceil(tiles / CUs)
    # launches of one tile per CU: the figures are per product (HIP events on the library
    # stream around its launches).
    flops = 2.0 * ntr * ntr * (d + k)
    t1, t2 = bf16_info["lra"]["t_product1"], bf16_info["lra"]["t_product2"]
    tr = KPROD_TRAFFIC
    if low:
        peak, kname = PEAK_BF16, ("kprod_wide<8,64,EM_F16,4,2> (fused RBF product, 256-row x "
                                  "256-column tiles, 64 x 256 per wave, one cooperative launch "
                                  "per product with tiles taken per XCD; K_Y Omega and K_Y W)")
    else:
        peak, kname = PEAK_FP32, "kprod_f32 (fused RBF product, K_Y Omega and K_Y W)"
    t_kp = 0.5 * (t1 + t2)
    achieved = flops / t_kp / 1e12
    same_config = args.n == 2_000_000 and k == 512 and args.precision == "bf16" and args.data == "gaussian"
    roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4),
                "traffic": (tr["fetch_bytes"] + tr["write_bytes"]
                            if tr["fetch_bytes"] is not None and same_config else None),
                "kernel": kname, "flops_per_product": flops, "product_s": round(t_kp, 4),
                "launches_per_product": (1 if not low or -(-ntr // 256) * -(-k // 256) >= 512
                                         else -(-(-(-ntr // 256) * -(-k // 256)) // 256)),
                "products": {"K_Y Omega": round(t1, 4), "K_Y W": round(t2, 4)},
                "traffic_source": tr["source"] if same_config else None}

    grad = None if args.no_grad else grad_form(ctx, torch)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(X, y, gamma, k, ntr, int(bf16_info["ipm"]["iters"]))

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(value * 1000, 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"bf16": "bf16/fp16", "fp8": "bf16/fp16/fp8-e4m3", "fp32": "fp32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"{'gaussian-mixture' if args.data == 'gaussian' else 'covtype-like'} "
                               f"n={args.n / 1e6:g}M ({ntr / 1e6:g}M train / "
                               f"{(args.n - ntr) / 1e6:g}M held-out), d={d}, k={k}, "
                               f"gamma={gamma:g}, C=32, stop: relative gap 1e-5",
                   "n_train": ntr, "d": d, "k": k, "gamma": gamma, "C": C, "seed": seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bf16": None,
        "fp32": fp32,
        "speedup_bf16_vs_fp32": (round(fp32["train_s"] / value, 2) if fp32 else None),
        "omega_gen_s": round(t_omega, 2),
        "zdz": {"form": "hybrid capacitance (error-budgeted split by w_i |z_i|^2): a dense BF16 "
                        "(or FP8) MFMA pass over the small-weight rows, compensated hi+lo bf16 "
                        "MFMA rows in the middle band, FP64 MFMA rows at the top",
                "bf16_tflops": round(bf16_info["ipm"]["syrk_bf16_tflops"], 1),
                "bf16_frac_of_peak": round(bf16_info["ipm"]["syrk_bf16_tflops"] / PEAK_BF16, 4),
                "bf16_row_frac": round(bf16_info["ipm"]["bf16_row_frac"], 4),
                "fp8_row_frac": round(bf16_info["ipm"]["fp8_row_frac"], 4),
                "x2_row_frac": round(bf16_info["ipm"]["x2_row_frac"], 4),
                "bf16_iters": bf16_info["ipm"]["syrk_bf16_iters"],
                "t_capacitance_s": round(bf16_info["ipm"]["t_syrk"], 3),
                "t_bf16_part_s": round(bf16_info["ipm"]["t_syrk_bf16"], 3),
                "fp32_form_fp64_tflops": (round(fp32["info"]["ipm"]["syrk_tflops"], 1) if fp32 else None),
                "capacitance_tflops": round(bf16_info["ipm"]["syrk_tflops"], 1),
                "capacitance_frac_of_bf16_peak": round(bf16_info["ipm"]["syrk_tflops"] / PEAK_BF16, 4),
                "capacitance_ms_per_iter": round(1e3 * bf16_info["ipm"]["t_syrk"]
                                                 / max(bf16_info["ipm"]["iters"], 1), 3),
                # the dense pass alone: every row through the lower 256 x 256 tiles
                # (diagonal tiles whole); "mfma" counts the MFMA work it issues, "syrk"
                # the SYRK convention n k (k + 1) over all n rows
                **dense_pass_figures(bf16_info["ipm"], ntr, k)},
        "grad": grad,
        "grad_timed": grad_timed,
        "fp64_ref": fp64_ref,
    }
    # the timed form under its own name; with BF16 timed, "fp32" holds the comparison run
    out[args.precision] = {"train_s": value, "accuracy": acc_bf16, "info": bf16_info}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
