"""Writes the committed golden fixtures under tests/golden/ (run in the build container).

Fixtures (small, numpy .npz / text):
  glibc_rand.npz   first 8192 raw glibc rand() values from the default (seed 1) state,
                   read through ctypes from the C library itself; and the first 8192
                   gaussrand() outputs recomputed here in numpy float32 from that raw
                   stream (the polar method of tensorsvm-train.cu:2621-2643; log is the
                   C library's logf, called through ctypes, as the reference's C++
                   float overload of log resolves to it).
  ipm_svc_*.npz    an independent QP solution for small low-rank SVM duals:
                   sklearn.svm.SVC(kernel='precomputed') (LIBSVM's SMO) on the Gram
                   matrix diag(y) Z Z^T diag(y); alpha and b = intercept_.
  model_rbf.txt    expected LIBSVM model text for a hand-made alpha (formats of
                   tensorsvm-train.cu:384-406), and model_linear.txt (340-350).
  tiny.libsvm      a small LIBSVM text file with its expected dense matrix (reader).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def glibc_fixture(count=8192):
    libc = ctypes.CDLL("libc.so.6")
    libm = ctypes.CDLL("libm.so.6")
    libm.logf.restype = ctypes.c_float
    libm.logf.argtypes = [ctypes.c_float]
    libc.srand(1)
    raw = np.array([libc.rand() for _ in range(count * 2)], dtype=np.int64)
    # gaussrand in float32 arithmetic
    out = []
    pos = 0
    f32 = np.float32
    rmax = f32(2147483647)  # RAND_MAX converted to float (= 2^31)
    while len(out) < count:
        while True:
            u1 = f32(raw[pos]) / rmax
            u2 = f32(raw[pos + 1]) / rmax
            pos += 2
            v1 = f32(2) * u1 - f32(1)
            v2 = f32(2) * u2 - f32(1)
            s = f32(v1 * v1) + f32(v2 * v2)
            if not (s >= f32(1) or s == f32(0)):
                break
        lg = f32(libm.logf(float(s)))
        q = f32(f32(-2) * lg) / s
        sq = np.sqrt(q, dtype=np.float32)
        out.append(f32(v1 * sq))
        out.append(f32(v2 * sq))
    gauss = np.array(out[:count], dtype=np.float32)
    np.savez(os.path.join(HERE, "glibc_rand.npz"), rand=raw[:count].astype(np.int32),
             gauss_bits=gauss.view(np.uint32))


def ipm_fixture(name, n, k, C, seed):
    from sklearn.svm import SVC
    rng = np.random.default_rng(seed)
    y = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    # low-rank factor of a labelled kernel: U = features with class signal
    U = rng.normal(size=(n, k)) * 0.6 + 0.4 * y[:, None] * rng.normal(size=(1, k))
    Z = y[:, None] * U                      # labelled factor (Z Z^T = diag(y) U U^T diag(y))
    K = U @ U.T                             # unlabelled Gram
    svc = SVC(kernel="precomputed", C=C, tol=1e-12, max_iter=10_000_000)
    svc.fit(K, y)
    alpha = np.zeros(n)
    alpha[svc.support_] = np.abs(svc.dual_coef_[0])
    np.savez(os.path.join(HERE, f"ipm_svc_{name}.npz"), Z=Z, y=y, C=C, alpha=alpha,
             b=float(svc.intercept_[0]))


def model_fixture():
    n, d = 6, 4
    alpha = np.array([0.0, 0.5, 1.0, 1e-7, 0.25, 1.0 - 1e-7])
    labels = np.array([1, -1, 1, -1, -1, 1], dtype=np.float32)
    inst = np.array([[0.5, 0, -1.25, 3],
                     [0, 0, 0, 0],
                     [1e-3, 2, 0, -0.5],
                     [1, 1, 1, 1],
                     [0, -7.125, 0, 0.0625],
                     [3.5, 0, 0, 0]], dtype=np.float32)
    C, gamma, b, pos, neg = 1.0, 0.05, -0.125, 3, 7
    nsv = int(np.sum(alpha > 1e-6))
    nfree = int(np.sum((alpha > 1e-6) & (alpha < C - 1e-6)))
    lines = ["svm_type c_svc", "kernel_type rbf", "gamma %.7f" % gamma, "nr_class 2",
             "total_sv %d" % nsv, "rho %f" % (-b), "label %d %d" % (pos, neg),
             "nr_sv %d %d" % (nfree, nsv - nfree), "SV"]
    for j in range(n):
        if not alpha[j] > 1e-6:
            continue
        s = "%7f " % (float(labels[j]) * alpha[j])
        for t in range(d):
            v = float(inst[j, t])
            if v != 0.0:
                s += "%d:%7f " % (t + 1, v)
        lines.append(s)
    with open(os.path.join(HERE, "model_rbf.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")
    w = np.array([0.1, -2.0, 1.0 / 3.0, 0.0])
    lin = ["solver_type L2R_L1LOSS_SVC_DUAL", "nr_class 2", "label %d %d" % (pos, neg),
           "nr_feature %d" % d, "bias 1", "w"] + ["%.16f " % v for v in w] + ["%.16f" % b]
    with open(os.path.join(HERE, "model_linear.txt"), "w") as f:
        f.write("\n".join(lin) + "\n")
    np.savez(os.path.join(HERE, "model_inputs.npz"), alpha=alpha, labels=labels, inst=inst,
             C=C, gamma=gamma, b=b, pos=pos, neg=neg, w=w)


def reader_fixture():
    text = ("+1 1:0.5 3:-1.25 4:3\n"
            "-1 2:7\n"
            "+1 1:1e-3 2:2 4:-0.5\n"
            "-1 5:1\n")
    with open(os.path.join(HERE, "tiny.libsvm"), "w") as f:
        f.write(text)
    dense = np.zeros((4, 5), dtype=np.float32)
    dense[0, [0, 2, 3]] = [0.5, -1.25, 3]
    dense[1, 1] = 7
    dense[2, [0, 1, 3]] = [1e-3, 2, -0.5]
    dense[3, 4] = 1
    np.savez(os.path.join(HERE, "tiny_libsvm.npz"), x=dense,
             labels=np.array([1, -1, 1, -1], dtype=np.float32))


if __name__ == "__main__":
    glibc_fixture()
    ipm_fixture("a", 240, 6, 1.0, 11)
    ipm_fixture("b", 300, 12, 32.0, 12)
    model_fixture()
    reader_fixture()
    print("golden fixtures written to", HERE)
