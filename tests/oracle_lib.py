"""ctypes binding of the CPU oracle (oracle/liboracle.so) for tests, smoke and bench.

The oracle is the checker only: product code never imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
D = ctypes.c_double


class IpmInfo(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("reason", ctypes.c_int), ("mu", D),
                ("normdx", D), ("normdy", D), ("pobj", D), ("dobj", D), ("gap_rel", D),
                ("condD", D), ("y", D)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_gaussrand_fill.argtypes = [P, I64]
        L.orc_glibc_rand.argtypes = [P, I64]
        L.orc_libsvm_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(P),
                                      ctypes.POINTER(P), ctypes.POINTER(I64), ctypes.POINTER(I64)]
        L.orc_free.argtypes = [P]
        L.orc_label_map.argtypes = [P, I64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_label_map.restype = I64
        L.orc_kernel_product.argtypes = [P, P, I64, P, P, I64, I64, P, I64, D, P]
        L.orc_kernel_product_f32.argtypes = [P, P, I64, P, P, I64, I64, P, I64, D, P]
        L.orc_householder_q.argtypes = [P, I64, I64]
        L.orc_sym_eig.argtypes = [P, I64, P, P]
        L.orc_lra.argtypes = [P, P, I64, I64, I64, D, P, P, P]
        L.orc_lra.restype = I64
        L.orc_lra_ex.argtypes = [P, P, I64, I64, I64, D, P, ctypes.c_int, P, P]
        L.orc_lra_ex.restype = I64
        L.orc_mpc.argtypes = [P, P, I64, I64, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, D,
                              P, P, P, ctypes.POINTER(IpmInfo)]
        L.orc_smw_solve.argtypes = [P, P, I64, I64, P]
        L.orc_newton_direction.argtypes = [P, P, I64, I64, D, P, P, P, P]
        L.orc_bias_rbf.argtypes = [P, P, P, I64, I64, D, P, P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(D)]
        L.orc_bias_rbf.restype = D
        L.orc_bias_linear.argtypes = [P, P, P, I64, I64, D, P]
        L.orc_bias_linear.restype = D
        L.orc_write_model_rbf.argtypes = [ctypes.c_char_p, P, P, P, I64, I64, D, D, D,
                                          ctypes.c_int, ctypes.c_int]
        L.orc_write_model_linear.argtypes = [ctypes.c_char_p, P, I64, D, ctypes.c_int, ctypes.c_int]
        L.orc_decision.argtypes = [P, P, P, I64, I64, D, D, P, I64, P]
        L.orc_fnorm_res.argtypes = [P, P, I64, I64, D, P, I64, ctypes.POINTER(D), ctypes.POINTER(D)]
        L.orc_grad.argtypes = [P, P, I64, I64, I64, P, P, D, D, P, P, P, I64, P, P, P]
        L.orc_grad.restype = ctypes.c_int
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_num_threads.argtypes = [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(P)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gaussrand(count):
    out = np.empty(count, dtype=np.float32)
    lib().orc_gaussrand_fill(_p(out), count)
    return out


def glibc_rand(count):
    out = np.empty(count, dtype=np.int32)
    lib().orc_glibc_rand(_p(out), count)
    return out


def libsvm_read(path, scale256=False):
    lab, inst, n, d = P(), P(), I64(), I64()
    rc = lib().orc_libsvm_read(path.encode(), int(scale256), ctypes.byref(lab), ctypes.byref(inst),
                               ctypes.byref(n), ctypes.byref(d))
    if rc != 0:
        raise IOError(f"oracle reader failed with {rc}")
    nn, dd = n.value, d.value
    labels = np.ctypeslib.as_array(ctypes.cast(lab, ctypes.POINTER(ctypes.c_float)), (max(nn, 1),))[:nn].copy()
    x = np.ctypeslib.as_array(ctypes.cast(inst, ctypes.POINTER(ctypes.c_float)), (max(nn * dd, 1),))[:nn * dd].copy()
    lib().orc_free(lab)
    lib().orc_free(inst)
    return labels, x.reshape(nn, dd)


def label_map(labels):
    labels = f32(labels).copy()
    pos, neg = ctypes.c_int(), ctypes.c_int()
    bad = lib().orc_label_map(_p(labels), labels.size, ctypes.byref(pos), ctypes.byref(neg))
    return labels, pos.value, neg.value, bad


def kernel_product(X1, y1, X2, y2, A, gamma):
    X1, y1, X2, y2, A = f32(X1), f32(y1), f32(X2), f32(y2), f64(A)
    m, d = X1.shape
    n = X2.shape[0]
    r = A.shape[1]
    W = np.empty((m, r), dtype=np.float64)
    lib().orc_kernel_product(_p(X1), _p(y1), m, _p(X2), _p(y2), n, d, _p(A), r, gamma, _p(W))
    return W


def kernel_product_f32(X1, y1, X2, y2, A, gamma):
    """The reference's FP32 arithmetic of the same product (float result)."""
    X1, y1, X2, y2, A = f32(X1), f32(y1), f32(X2), f32(y2), f32(A)
    m, d = X1.shape
    n = X2.shape[0]
    r = A.shape[1]
    W = np.empty((m, r), dtype=np.float32)
    lib().orc_kernel_product_f32(_p(X1), _p(y1), m, _p(X2), _p(y2), n, d, _p(A), r, gamma, _p(W))
    return W


def householder_q(W):
    W = f64(W).copy()
    n, k = W.shape
    lib().orc_householder_q(_p(W), n, k)
    return W


def sym_eig(A):
    A = f64(A)
    k = A.shape[0]
    w = np.empty(k)
    V = np.empty((k, k))
    lib().orc_sym_eig(_p(A), k, _p(w), _p(V))
    return w, V


LRA_FP64 = 0
LRA_REF_FP32 = 1


def lra(X, y, k, gamma, omega, mode=LRA_FP64):
    """Rank-k factor; mode LRA_REF_FP32 follows the reference's FP32 storage of W, Q, C,
    X and U (the plain method's precision path)."""
    X, y, omega = f32(X), f32(y), f32(omega)
    n, d = X.shape
    U = np.empty((n, k))
    w = np.empty(k)
    realk = lib().orc_lra_ex(_p(X), _p(y), n, d, k, gamma, _p(omega), mode, _p(U), _p(w))
    return U, w, realk


def mpc(Z, a, C, kernel_type=2, maxit=100, stop="ref", tol=1e-5):
    Z, a = f64(Z), f64(a)
    n, k = Z.shape
    X, Xi, S = np.empty(n), np.empty(n), np.empty(n)
    info = IpmInfo()
    lib().orc_mpc(_p(Z), _p(a), n, k, C, kernel_type, maxit, 0 if stop == "ref" else 1, tol,
                  _p(X), _p(Xi), _p(S), ctypes.byref(info))
    return X, Xi, S, info


def smw_solve(Z, Dv, b):
    Z, Dv = f64(Z), f64(Dv)
    b = f64(b).copy()
    n, k = Z.shape
    lib().orc_smw_solve(_p(Z), _p(Dv), n, k, _p(b))
    return b


def newton_direction(Z, a, C, X, S, Xi, r):
    Z, a, X, S, Xi = f64(Z), f64(a), f64(X), f64(S), f64(Xi)
    r = f64(r).copy()
    n, k = Z.shape
    lib().orc_newton_direction(_p(Z), _p(a), n, k, C, _p(X), _p(S), _p(Xi), _p(r))
    return r


def bias_rbf(alpha, labels, U, C, S=None, Xi=None):
    """Reference free set without slacks; with the IPM's final S and Xi, the trainer's
    slack-confirmed free set."""
    alpha, labels, U = f64(alpha), f32(labels), f64(U)
    n, k = U.shape
    cnt, sd = ctypes.c_int(), D()
    S = None if S is None else f64(S)
    Xi = None if Xi is None else f64(Xi)
    ps = None if S is None else _p(S)
    px = None if Xi is None else _p(Xi)
    b = lib().orc_bias_rbf(_p(alpha), _p(labels), _p(U), n, k, C, ps, px, ctypes.byref(cnt),
                           ctypes.byref(sd))
    return b, cnt.value, sd.value


def bias_linear(alpha, labels, inst, C):
    alpha, labels, inst = f64(alpha), f32(labels), f32(inst)
    n, d = inst.shape
    w = np.empty(d)
    b = lib().orc_bias_linear(_p(alpha), _p(labels), _p(inst), n, d, C, _p(w))
    return b, w


def write_model_rbf(path, alpha, labels, inst, C, gamma, b, pos, neg):
    alpha, labels, inst = f64(alpha), f32(labels), f32(inst)
    n, d = inst.shape
    return lib().orc_write_model_rbf(path.encode(), _p(alpha), _p(labels), _p(inst), n, d, C,
                                     gamma, b, pos, neg)


def write_model_linear(path, w, b, pos, neg):
    w = f64(w)
    return lib().orc_write_model_linear(path.encode(), _p(w), w.size, b, pos, neg)


def decision(alpha, labels, inst, gamma, b, test):
    alpha, labels, inst, test = f64(alpha), f32(labels), f32(inst), f32(test)
    n, d = inst.shape
    f = np.empty(test.shape[0])
    lib().orc_decision(_p(alpha), _p(labels), _p(inst), n, d, gamma, b, _p(test), test.shape[0], _p(f))
    return f


def fnorm_res(X, y, gamma, U):
    X, y, U = f32(X), f32(y), f64(U)
    n, d = X.shape
    a, b = D(), D()
    lib().orc_fnorm_res(_p(X), _p(y), n, d, gamma, _p(U), U.shape[1], ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def grad(X, y, k, omega, alpha, b, log_gamma, log_scales, Xt=None, yt=None, want_grad=True):
    """FP64 gradients of the trained dual D and the held-out mean smoothed hinge L with
    respect to (log gamma, log_scales); returns (D, L, gD, gL, status)."""
    X, y = f32(X), f32(y)
    n, d = X.shape
    om = f32(omega).reshape(-1)
    alpha = f64(alpha)
    sc = f64(log_scales)
    if Xt is None:
        Xt = np.zeros((0, d), np.float32)
        yt = np.zeros(0, np.float32)
    Xt, yt = f32(Xt), f32(yt)
    nt = Xt.shape[0]
    vals = np.zeros(2)
    gD = np.zeros(d + 1)
    gL = np.zeros(d + 1)
    st = lib().orc_grad(_p(X), _p(y), n, d, k, _p(om), _p(alpha), b, log_gamma, _p(sc),
                        _p(Xt) if nt else None, _p(yt) if nt else None, nt, _p(vals),
                        _p(gD) if want_grad else None, _p(gL) if want_grad else None)
    return vals[0], vals[1], gD, gL, st

