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
  roofline      the fused kernel-product kernel (>98% of training flops): algorithmic
                flops 2 m n (d + k) per launch / its launch time (HIP events on the
                library's stream), against the BF16 dense MFMA peak
  cpu_baseline  the CPU oracle's kernel product (FP64, all host threads) on a bounded row
                sample, extrapolated to the two full products of the factor construction
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=2_000_000, help="rows generated (80%% train)")
    ap.add_argument("--k", type=int, default=512)
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-rows", type=int, default=384)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the trainer is "
              "single-GPU and multi-GPU sharding is not part of this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import tensorsvm_amd as T
    from datagen import gaussian_mixture

    d, gamma, C, k = 18, 0.05, 32.0, args.k
    seed = 1002
    t = time.time()
    Xall, yall = gaussian_mixture(args.n, d, seed)
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
                                  stop=T.capi.STOP_GAP, tol=1e-5, maxit=300, refine=1)
        t0 = time.perf_counter()
        info = ctx.train_dev(dX, dy, ntr, d, prm, dOm, dal)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, info

    def accuracy(info):
        alpha = dal.cpu().numpy()
        f = ctx.decision_function(X, y, alpha, gamma, info.b, Xte)
        return float(np.mean(np.sign(f) == yte)), alpha

    for w in range(args.warmup):
        dt, _ = run(T.PREC_BF16)
        log(f"warmup {w}: {dt:.2f}s")
    times, info = [], None
    for s in range(args.steps):
        dt, info = run(T.PREC_BF16)
        times.append(dt)
        log(f"step {s}: bf16 train {dt:.2f}s (lra {info.lra.t_total:.2f}s, ipm {info.ipm.t_total:.2f}s, "
            f"{info.ipm.iters} it, reason {info.ipm.reason}, gap {info.ipm.gap_rel:.2e})")
    acc_bf16, alpha_bf16 = accuracy(info)
    bf16_info = info.as_dict()
    value = float(np.mean(times))

    fp32 = None
    if not args.no_fp32:
        dt32, info32 = run(T.PREC_FP32)
        acc32, alpha32 = accuracy(info32)
        fp32 = {"train_s": dt32, "accuracy": acc32, "info": info32.as_dict(),
                "alpha_max_abs_diff_vs_bf16": float(np.abs(alpha32 - alpha_bf16).max())}
        log(f"fp32 train {dt32:.2f}s (lra {info32.lra.t_total:.2f}s), accuracy {acc32:.4f}")

    # roofline of the fused kernel-product kernel (second product of the last BF16 step)
    flops = 2.0 * ntr * ntr * (d + k)
    t_kp = bf16_info["lra"]["t_product2"]
    achieved = flops / t_kp / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_BF16, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_BF16, 4), "traffic": None,
                "kernel": "kprod_bf16 (fused RBF kernel product)",
                "flops_per_launch": flops}

    cpu = None
    if not args.no_cpu:
        import oracle_lib as O
        m = args.cpu_rows
        A = np.random.default_rng(3).standard_normal((ntr, k))
        t0 = time.perf_counter()
        O.kernel_product(X[:m], y[:m], X, y, A, gamma)
        tc = time.perf_counter() - t0
        est = tc * (ntr / m) * 2.0
        cpu = {"value": round(est, 1), "unit": "s", "cores": O.lib().orc_num_threads(), "kind": "port",
               "sample": f"oracle kernel product (FP64) of {m} rows x {ntr} columns x k={k} took "
                         f"{tc:.1f}s; extrapolated x{ntr / m:.0f} rows x 2 products (the factor "
                         "construction's >98% of flops; IPM not included)"}
        log(f"cpu baseline: sample {tc:.1f}s -> {est:.0f}s")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(value * 1000, 1), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "gaussian-mixture n=2M (1.6M train / 0.4M held-out), d=18, k=512, "
                               "gamma=0.05, C=32, stop: relative gap 1e-5",
                   "n_train": ntr, "d": d, "k": k, "gamma": gamma, "C": C, "seed": seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bf16": {"train_s": value, "accuracy": acc_bf16, "info": bf16_info},
        "fp32": fp32,
        "speedup_bf16_vs_fp32": (round(fp32["train_s"] / value, 2) if fp32 else None),
        "omega_gen_s": round(t_omega, 2),
        "zdz": {"form": "FP64 MFMA capacitance (both forms)",
                "tflops": bf16_info["ipm"]["syrk_tflops"],
                "frac_of_fp64_peak": round(bf16_info["ipm"]["syrk_tflops"] / PEAK_FP64, 4)},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
