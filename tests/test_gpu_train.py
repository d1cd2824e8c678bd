"""End-to-end training (factor + IPM + bias) and the command line, against the oracle's
plain method on the same seeded inputs and Omega."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
from datagen import gaussian_mixture

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _split(n, d, seed):
    X, y = gaussian_mixture(n, d, seed)
    ntr = int(0.8 * n)
    return X[:ntr], y[:ntr], X[ntr:], y[ntr:]


def _oracle_train(X, y, k, gamma, C, om, mode=O.LRA_FP64):
    """The oracle's plain method: the factor (LRA 1941-2249; mode LRA_REF_FP32 in the
    reference's FP32 arithmetic and storage), the FP64 IPM (mpc 1507-1840) and the bias
    over the first <= 50 free SVs (writemodel 353-372)."""
    U, w, realk = O.lra(X, y, k, gamma, om, mode)
    al, xi, s, info = O.mpc(U, y.astype(np.float64), C)
    b, cnt, sd = O.bias_rbf(al, y, U, C)
    return al, b, info, U


def _free_set(al, C):
    """The reference's bias samples: the first <= 50 indices with 1e-6 < alpha < C - 1e-6."""
    return np.flatnonzero((al > 1e-6) & (al < C - 1e-6))[:50]


def check_train_parity(name, ctx, X, y, Xt, yt, k, gamma, C, om, al, info):
    """Parity of an FP32-form training run (alpha, info) with the oracle's plain method.

    * The solver on the GPU's own factor (tsvm_lra with the same Omega) against the
      oracle's FP64 IPM and bias on that factor: alpha within 1e-5 C, the same
      support-vector set apart from multipliers within 1e-5 of 0 or C, the bias within
      1e-4 on the same free-SV samples.
    * End to end against the oracle in FP64 arithmetic: the QP is near-degenerate, so
      FP32 rounding of the factor moves alpha well past its own size; the bar is 1e-4 C or
      three times the distance between the oracle in the reference's FP32 arithmetic and
      the FP64 oracle, whichever is larger -- the GPU form is as close to exact arithmetic
      as the reference's own arithmetic is.  The primal objective to 1e-6 and held-out
      accuracy within 0.1 point (one row when fewer than 1000)."""
    import tensorsvm_amd as T
    from conftest import parity_log
    om = om.reshape(-1, k)
    alo, bo, io, Uo = _oracle_train(X, y, k, gamma, C, om, O.LRA_FP64)
    alr, br, ir, Ur = _oracle_train(X, y, k, gamma, C, om, O.LRA_REF_FP32)
    spread = np.abs(alr - alo).max() / C
    Ug, _ = ctx.lra(X, y, k, gamma, T.PREC_FP32, omega=om)
    als, _, _, isv = O.mpc(Ug.astype(np.float64), y.astype(np.float64), C)
    bs, _, _ = O.bias_rbf(als, y, Ug.astype(np.float64), C)
    d_same = np.abs(al - als).max() / C
    d_e2e = np.abs(al - alo).max() / C
    sg, ss = _free_set(al, C), _free_set(als, C)
    same = len(sg) == len(ss) and np.array_equal(sg, ss)
    fg = ctx.decision_function(X, y, al, gamma, info.b, Xt)
    fo = O.decision(alo, y, X, gamma, bo, Xt)
    acc_g, acc_o = np.mean(np.sign(fg) == yt), np.mean(np.sign(fo) == yt)
    parity_log(name, dalpha_same_factor=d_same, dalpha_e2e=d_e2e, ref_fp32_spread=spread,
               db_same_factor=abs(info.b - bs), same_bias_samples=bool(same),
               db_e2e=abs(info.b - bo), dpobj_rel=abs(info.ipm.pobj - io.pobj) / (1 + abs(io.pobj)),
               acc_gpu=acc_g, acc_oracle=acc_o, iters_gpu=info.ipm.iters, iters_oracle=io.iters)
    assert isv.reason == 0
    assert d_same <= 1e-5, d_same
    amb = (als < 1e-5) | (als > C - 1e-5)
    assert np.all(((al > 1e-6) == (als > 1e-6)) | amb)
    if same:
        assert abs(info.b - bs) <= 1e-4, (info.b, bs)
    assert d_e2e <= max(1e-4, 3 * spread), (d_e2e, spread)
    assert abs(info.ipm.pobj - io.pobj) <= 1e-6 * (1 + abs(io.pobj))
    assert abs(acc_g - acc_o) <= max(0.001, 1.0 / len(yt)) + 1e-12, (acc_g, acc_o)
    return fg


# (25000, 32): the plain method's small configuration, 20,000 training rows, d = 18,
# gamma = 0.05, C = 32, k = 32 (the oracle takes a few seconds on the host cores)
@pytest.mark.parametrize("n,k", [(3000, 32), (4000, 64), (25000, 32)])
def test_train_fp32_matches_oracle(ctx, n, k):
    import tensorsvm_amd as T
    gamma, C = 0.05, 32.0
    X, y, Xt, yt = _split(n, 18, 1000 + n)
    om = O.gaussrand(X.shape[0] * k)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=gamma, precision=T.PREC_FP32, omega=om)
    assert info.ipm.reason == 0
    assert info.b_slack_rule == 0  # the reference's free set under the reference stop
    fg = check_train_parity(f"train_fp32[{n}-{k}]", ctx, X, y, Xt, yt, k, gamma, C, om, al, info)
    # the GPU decision function (FP32 products) against FP64 on the same multipliers, to
    # the FP32 accumulation error of sum_j |alpha_j| K(x_t, x_j)
    scale = O.decision(np.abs(al), np.ones_like(y), X, gamma, 0.0, Xt)
    err = np.abs(fg - O.decision(al, y, X, gamma, info.b, Xt))
    assert (err <= 1e-4 + 1e-5 * scale).all(), (err.max(), scale.max())


@pytest.mark.parametrize("form", ["bf16", "split"])
def test_train_bf16_same_accuracy(ctx, form):
    import tensorsvm_amd as T
    prec = T.PREC_BF16 if form == "bf16" else T.PREC_FP16X2
    gamma, C, k = 0.05, 32.0, 64
    X, y, Xt, yt = _split(6000, 18, 77)
    om = O.gaussrand(X.shape[0] * k)
    al32, _, i32 = ctx.train(X, y, k=k, C=C, gamma=gamma, precision=T.PREC_FP32, omega=om,
                             stop=T.capi.STOP_GAP, tol=1e-5)
    al16, _, i16 = ctx.train(X, y, k=k, C=C, gamma=gamma, precision=prec, omega=om,
                             stop=T.capi.STOP_GAP, tol=1e-5)
    assert i16.ipm.reason == 0 and i32.ipm.reason == 0
    assert abs(i16.ipm.pobj - i32.ipm.pobj) <= 5e-3 * abs(i32.ipm.pobj)
    a32 = np.mean(np.sign(ctx.decision_function(X, y, al32, gamma, i32.b, Xt)) == yt)
    a16 = np.mean(np.sign(ctx.decision_function(X, y, al16, gamma, i16.b, Xt)) == yt)
    assert abs(a32 - a16) <= 0.01


def _write_libsvm(path, X, y):
    with open(path, "w") as f:
        for xi, yi in zip(X, y):
            f.write(("+1" if yi > 0 else "-1") + " " +
                    " ".join(f"{j + 1}:{v:.7g}" for j, v in enumerate(xi) if v != 0) + "\n")


def _read_model(path):
    head, svs = {}, []
    with open(path) as f:
        lines = f.read().splitlines()
    i = lines.index("SV")
    for ln in lines[:i]:
        kv = ln.split(" ", 1)
        head[kv[0]] = kv[1]
    for ln in lines[i + 1:]:
        parts = ln.split()
        svs.append(float(parts[0]))
    return head, np.array(svs)


def test_cli_rbf_model_file(tmp_path):
    X, y = gaussian_mixture(2000, 18, 5)
    tr = str(tmp_path / "train.libsvm")
    _write_libsvm(tr, X, y)
    model = str(tmp_path / "out.model")
    exe = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-train")
    r = subprocess.run([exe, "-k", "32", "-c", "32", "-g", "0.05", "-t", "2", "-analysis", tr, model],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "found labels:" in r.stdout and "#BSV" in r.stdout
    assert "fnorm(K)=" in r.stdout and "fnorm(K-U*U=" in r.stdout
    head, coef = _read_model(model)
    assert head["svm_type"] == "c_svc" and head["kernel_type"] == "rbf"
    assert head["gamma"] == "0.0500000" and head["nr_class"] == "2"
    # the same data through the oracle (labels as read, Omega = the gaussrand stream)
    lab, xs = O.libsvm_read(tr)
    lab, pos, neg, bad = O.label_map(lab)
    assert head["label"] == f"{pos} {neg}"
    om = O.gaussrand(xs.shape[0] * 32).reshape(-1, 32)
    alo, bo, io, U = _oracle_train(xs, lab, 32, 0.05, 32.0, om)
    nsv = int(np.sum(alo > 1e-6))
    assert abs(int(head["total_sv"]) - nsv) <= max(2, nsv // 200)
    assert abs(float(head["rho"]) + bo) <= 5e-3 * (1 + abs(bo))
    assert np.all(np.abs(coef) <= 32.0 + 1e-6)


@pytest.mark.parametrize("prec", ["bf16", "split"])
def test_cli_low_precision_forms(tmp_path, prec):
    """-precision bf16 / split through the command line: a model file whose held-out
    accuracy (through tensorsvm-predict) matches the FP32 form's."""
    X, y = gaussian_mixture(3000, 18, 21)
    tr, te = str(tmp_path / "tr.libsvm"), str(tmp_path / "te.libsvm")
    _write_libsvm(tr, X[:2400], y[:2400])
    _write_libsvm(te, X[2400:], y[2400:])
    exe = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-train")
    pred = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-predict")
    acc = {}
    for form in ("fp32", prec):
        model = str(tmp_path / f"{form}.model")
        r = subprocess.run([exe, "-k", "48", "-c", "32", "-g", "0.05", "-t", "2", "-precision", form,
                            "-stop", "gap", tr, model], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        out = str(tmp_path / f"{form}.out")
        r = subprocess.run([pred, te, model, out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        acc[form] = float(r.stdout.split("Accuracy = ")[1].split("%")[0])
    assert abs(acc[prec] - acc["fp32"]) <= 1.0, acc


def test_train_progress_lines(tmp_path):
    """progress = 1: one stderr line per kernel-product slab and per 10 IPM iterations, and
    the same multipliers as without."""
    script = tmp_path / "p.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
import numpy as np
import tensorsvm_amd as T
import oracle_lib as O
from datagen import gaussian_mixture
X, y = gaussian_mixture(3000, 18, 3)
om = O.gaussrand(3000 * 32)
c = T.Context(0)
a0, _, i0 = c.train(X, y, k=32, C=32.0, gamma=0.05, precision=T.PREC_BF16, omega=om,
                    stop=T.capi.STOP_GAP, tol=1e-5)
a1, _, i1 = c.train(X, y, k=32, C=32.0, gamma=0.05, precision=T.PREC_BF16, omega=om,
                    stop=T.capi.STOP_GAP, tol=1e-5, progress=1)
print("same", bool(np.array_equal(a0, a1)), i0.ipm.iters, i1.ipm.iters)
""")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "same True" in r.stdout, r.stdout
    assert "factor: K_Y Omega: rows 3000 / 3000" in r.stderr, r.stderr
    assert "factor: K_Y W: rows 3000 / 3000" in r.stderr
    assert "ipm: iteration 0, relative gap" in r.stderr


def test_cli_linear_model_file(tmp_path):
    rng = np.random.default_rng(8)
    n, d = 1500, 6
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    X = (rng.normal(size=(n, d)) + 0.7 * y[:, None]).astype(np.float32)
    tr = str(tmp_path / "lin.libsvm")
    _write_libsvm(tr, X, y)
    exe = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-train")
    r = subprocess.run([exe, "-t", "0", "-c", "1", tr], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = open(tr + ".model").read().splitlines()
    assert lines[0] == "solver_type L2R_L1LOSS_SVC_DUAL" and lines[5] == "w"
    w = np.array([float(v) for v in lines[6:6 + d]])
    lab, xs = O.libsvm_read(tr)
    lab, pos, neg, bad = O.label_map(lab)
    Z = lab[:, None].astype(np.float64) * xs
    al, xi, s, info = O.mpc(Z, lab, 1.0, kernel_type=0)
    bo, wo = O.bias_linear(al, lab, xs, 1.0)
    assert np.allclose(w, wo, atol=1e-4 * (1 + np.abs(wo).max()))
    assert abs(float(lines[6 + d]) - bo) <= 1e-3 * (1 + abs(bo))


def test_predict_from_written_rbf_model(ctx, tmp_path):
    """The predictor on the written model file reproduces the in-memory decision function
    (up to the file's 6-decimal coefficients and features) and the oracle's FP64 one; the
    command line prints the matching accuracy."""
    import tensorsvm_amd as T
    gamma, C, k = 0.05, 32.0, 32
    X, y, Xt, yt = _split(3000, 18, 41)
    al, _, info = ctx.train(X, y, k=k, C=C, gamma=gamma, precision=T.PREC_FP32,
                            omega=O.gaussrand(X.shape[0] * k))
    path = str(tmp_path / "m.model")
    T.capi.write_model_rbf(path, al, y, X, C, gamma, info.b, 1, -1)
    md = T.Model(path)
    assert md.nsv == int((al > 1e-6).sum())
    f_file, pred = ctx.predict(md, Xt)
    f_mem = ctx.decision_function(X, y, al, gamma, info.b, Xt)
    scale = np.abs(f_mem).max()
    assert np.abs(f_file - f_mem).max() <= 1e-4 * scale + 1e-5
    # FP64 oracle decision on the model's own (rounded) coefficients is the same function
    f_orc = O.decision(al, y, X, gamma, info.b, Xt)
    assert np.abs(f_file - f_orc).max() <= 1e-4 * scale + 1e-5
    sure = np.abs(f_mem) > 1e-3 * scale
    assert np.all(pred[sure] == np.where(f_mem[sure] > 0, 1.0, -1.0))
    te = str(tmp_path / "test.libsvm")
    _write_libsvm(te, Xt, yt)
    out = str(tmp_path / "pred.txt")
    exe = os.path.join(ROOT, "tensorsvm_amd", "tensorsvm-predict")
    r = subprocess.run([exe, te, path, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    p2 = np.array([float(v) for v in open(out).read().split()])
    acc = np.mean(p2 == yt)
    assert f"Accuracy = {100 * acc:g}%" in r.stdout, r.stdout
    assert np.mean(p2 == pred) >= 0.999


def test_predict_from_written_linear_model(ctx, tmp_path):
    import tensorsvm_amd as T
    rng = np.random.default_rng(9)
    n, d = 2000, 7
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0).astype(np.float32)
    X = (rng.normal(size=(n, d)) + 0.6 * y[:, None]).astype(np.float32)
    al, w, info = ctx.train(X, y, k=d, C=1.0, kernel_type=0)
    path = str(tmp_path / "l.model")
    T.capi.write_model_linear(path, w, info.b, 1, -1)
    md = T.Model(path)
    f, pred = ctx.predict(md, X)
    ref = X.astype(np.float64) @ w + info.b
    assert np.abs(f - ref).max() <= 1e-12 * (1 + np.abs(ref).max()) + 1e-9
    assert np.all(pred == np.where(ref > 0, 1.0, -1.0))

