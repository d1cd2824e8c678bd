"""ctypes binding of libtsvm's C-ABI (include/tsvm.h).

Host arrays are numpy; device arrays (for the ``*_dev`` entry points) are any object
exposing an integer device address (torch tensors' ``data_ptr()``).  Every call checks
the returned status and raises TsvmError with the library's last error text.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PREC_FP32 = 0
PREC_BF16 = 1
PREC_FP16X2 = 2  # kernel values as two fp16 terms; as a training form: BF16 with a two-term sketch
PREC_FP16 = 3    # kernel product only: kernel values and right-hand side in fp16
PREC_FP16R2 = 4  # kernel product only: right-hand side as two fp16 terms
PREC_FP64 = 5    # FP64 products (the reference of the other forms); as a training form:
                 # both factor products in FP64, the rest as the FP32 form
PREC_FP8 = 6     # training / IPM form: BF16 with an FP8 e4m3 capacitance tier

STOP_REF = 0
STOP_GAP = 1

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_INT = ctypes.c_int
_LIB = None


class TsvmError(RuntimeError):
    pass


def library_path():
    return os.path.join(HERE, "libtsvm.so")


def load():
    """Load libtsvm.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    # One HIP runtime per process: PyTorch ships its own libamdhip64 under the same
    # SONAME as ROCm's, and whichever is loaded first serves both.  Loading torch's first
    # lets device tensors and libtsvm share it (the other order leaves torch unable to
    # initialise its devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(path):
        raise TsvmError(f"{path} is missing: build it with `make -C tensorsvm_amd` "
                        "(or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    L.tsvm_create.argtypes = [_INT, ctypes.POINTER(_P)]
    L.tsvm_destroy.argtypes = [_P]
    L.tsvm_destroy.restype = None
    L.tsvm_last_error.argtypes = [_P]
    L.tsvm_last_error.restype = ctypes.c_char_p
    L.tsvm_sync.argtypes = [_P]
    L.tsvm_gaussrand_fill.argtypes = [_P, _I64, _INT]
    kp = [_P, _P, _P, _I64, _P, _P, _I64, _I64, _P, _I64, _D, _INT, _P]
    L.tsvm_kernel_product.argtypes = kp
    L.tsvm_kernel_product_dev.argtypes = kp
    L.tsvm_kernel_product_noalloc_dev.argtypes = kp
    _bind_optional(L)
    _LIB = L
    return L


def _bind_optional(L):
    """Bind entry points added beside the kernel product (same library)."""
    for name, args, res in _EXTRA:
        fn = getattr(L, name, None)
        if fn is not None:
            fn.argtypes = args
            if res == 0:
                fn.restype = None
            elif res is not None:
                fn.restype = res


class LraInfo(ctypes.Structure):
    _fields_ = [("t_product1", _D), ("t_orth", _D), ("t_product2", _D), ("t_core", _D),
                ("t_factor", _D), ("t_total", _D), ("realk", _I64), ("eig_max", _D),
                ("eig_min", _D), ("orth_passes", _INT)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class IpmParams(ctypes.Structure):
    _fields_ = [("C", _D), ("kernel_type", _INT), ("maxit", _INT), ("stop_mode", _INT),
                ("tol", _D), ("precision", _INT), ("refine", _INT), ("verbose", _INT),
                ("init", _INT)]


class IpmInfo(ctypes.Structure):
    _fields_ = [("iters", _INT), ("reason", _INT), ("mu", _D), ("normdx", _D), ("normdy", _D),
                ("pobj", _D), ("dobj", _D), ("gap_rel", _D), ("condD", _D), ("y", _D),
                ("t_total", _D), ("t_syrk", _D), ("syrk_tflops", _D), ("refine_res", _D),
                ("t_syrk_bf16", _D), ("syrk_bf16_tflops", _D), ("bf16_row_frac", _D),
                ("syrk_bf16_iters", _INT), ("fp8_row_frac", _D), ("x2_row_frac", _D),
                ("t_syrk_dense", _D), ("syrk_dense_mfma_tflops", _D)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SyrkInfo(ctypes.Structure):
    _fields_ = [("n_fp64_rows", _I64), ("bf16_bound", _D), ("t_bf16", _D), ("bf16_tflops", _D),
                ("n_fp8_rows", _I64), ("fp8_bound", _D), ("n_x2_rows", _I64), ("x2_bound", _D),
                ("t_dense", _D), ("dense_mfma_tflops", _D)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GradInfo(ctypes.Structure):
    _fields_ = [("t_forward", _D), ("t_backward", _D), ("backward_flops", _D),
                ("backward_tflops", _D), ("t_total", _D), ("realk", _I64), ("orth_passes", _INT),
                ("t_backward_heldout", _D)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def ipm_params(C, kernel_type=2, maxit=100, stop=STOP_REF, tol=1e-5, precision=PREC_FP32,
               refine=-1, verbose=0, init=0):
    return IpmParams(C, kernel_type, maxit, stop, tol, precision, refine, verbose, init)


class TrainParams(ctypes.Structure):
    _fields_ = [("k", _I64), ("C", _D), ("gamma", _D), ("kernel_type", _INT), ("precision", _INT),
                ("stop_mode", _INT), ("tol", _D), ("maxit", _INT), ("refine", _INT),
                ("verbose", _INT), ("init", _INT), ("analysis", _INT), ("progress", _INT)]


class TrainInfo(ctypes.Structure):
    _fields_ = [("lra", LraInfo), ("ipm", IpmInfo), ("b", _D), ("b_std", _D), ("b_samples", _INT),
                ("b_slack_rule", _INT), ("nsv", _I64), ("nfree", _I64), ("t_bias", _D), ("t_total", _D),
                ("fnorm", _D), ("fnorm_res", _D), ("t_analysis", _D)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("lra", "ipm")}
        d["lra"] = self.lra.as_dict()
        d["ipm"] = self.ipm.as_dict()
        return d


def train_params(k=50, C=1.0, gamma=0.0, kernel_type=2, precision=PREC_FP32, stop=STOP_REF,
                 tol=1e-5, maxit=100, refine=-1, verbose=0, init=0, analysis=0, progress=0):
    return TrainParams(k, C, gamma, kernel_type, precision, stop, tol, maxit, refine, verbose, init,
                       analysis, progress)


_EXTRA = [
    ("tsvm_train", [_P, _P, _P, _I64, _I64, ctypes.POINTER(TrainParams), _P, _P, _P,
                    ctypes.POINTER(TrainInfo)], None),
    ("tsvm_train_dev", [_P, _P, _P, _I64, _I64, ctypes.POINTER(TrainParams), _P, _P,
                        ctypes.POINTER(TrainInfo)], None),
    ("tsvm_decision_function", [_P, _P, _P, _P, _I64, _I64, _D, _D, _P, _I64, _P], None),
    ("tsvm_read_libsvm", [ctypes.c_char_p, _INT, ctypes.POINTER(_P), ctypes.POINTER(_P),
                          ctypes.POINTER(_I64), ctypes.POINTER(_I64)], None),
    ("tsvm_write_libsvm", [ctypes.c_char_p, _P, _P, _I64, _I64], None),
    ("tsvm_free", [_P], ctypes.c_int),
    ("tsvm_set_progress", [_P, _INT], 0),
    ("tsvm_map_labels", [_P, _I64, ctypes.POINTER(_INT), ctypes.POINTER(_INT)], _I64),
    ("tsvm_write_model_rbf", [ctypes.c_char_p, _P, _P, _P, _I64, _I64, _D, _D, _D, _INT, _INT], None),
    ("tsvm_write_model_linear", [ctypes.c_char_p, _P, _I64, _D, _INT, _INT], None),
    ("tsvm_ipm", [_P, _P, _P, _I64, _I64, ctypes.POINTER(IpmParams), _P, _P, _P,
                  ctypes.POINTER(IpmInfo)], None),
    ("tsvm_ipm_dev", [_P, _P, _P, _I64, _I64, ctypes.POINTER(IpmParams), _P, _P, _P,
                      ctypes.POINTER(IpmInfo)], None),
    ("tsvm_lra", [_P, _P, _P, _I64, _I64, _I64, _D, _INT, _P, _P, ctypes.POINTER(LraInfo)], None),
    ("tsvm_lra_dev", [_P, _P, _P, _I64, _I64, _I64, _D, _INT, _P, _P, ctypes.POINTER(LraInfo)], None),
    ("tsvm_lra_f64", [_P, _P, _I64, _P, _I64, _I64, _I64, _D, _INT, _P, _P, _I64, ctypes.POINTER(_D),
                      ctypes.POINTER(LraInfo)], None),
    ("tsvm_mpc", [_P, _P, _I64, _P, _D, _P, _P, _I64, _I64, ctypes.POINTER(IpmParams),
                  ctypes.POINTER(IpmInfo)], None),
    ("tsvm_weighted_syrk_dev", [_P, _P, _I64, _I64, _P, _INT, _D, _P, ctypes.POINTER(SyrkInfo)], None),
    ("tsvm_read_model", [ctypes.c_char_p, ctypes.POINTER(_P)], None),
    ("tsvm_model_free", [_P], 0),
    ("tsvm_model_get", [_P, ctypes.POINTER(_INT), ctypes.POINTER(_D), ctypes.POINTER(_D),
                        ctypes.POINTER(_INT), ctypes.POINTER(_I64), ctypes.POINTER(_I64)], None),
    ("tsvm_predict", [_P, _P, _P, _I64, _I64, _P, _P], None),
    ("tsvm_fnorm_res_dev", [_P, _P, _P, _I64, _I64, _D, _P, _I64, ctypes.POINTER(_D),
                            ctypes.POINTER(_D)], None),
    ("tsvm_grad_dev", [_P, _P, _P, _I64, _I64, _P, _D, _I64, _D, _P, _P, _P, _P, _I64, _INT, _P, _P,
                       _P, ctypes.POINTER(GradInfo)], None),
]



class Model:
    """A model file written by the trainer (LIBSVM RBF or LIBLINEAR text), parsed by
    tsvm_read_model; scored with Context.predict."""

    def __init__(self, path):
        L = load()
        self._lib = L
        h = _P()
        st = L.tsvm_read_model(os.fsencode(path), ctypes.byref(h))
        if st != 0:
            raise TsvmError(f"tsvm_read_model({path}) failed with status {st}")
        self.h = h
        kt, g, rho = _INT(), _D(), _D()
        lab = (_INT * 2)()
        nsv, d = _I64(), _I64()
        L.tsvm_model_get(h, ctypes.byref(kt), ctypes.byref(g), ctypes.byref(rho), lab,
                         ctypes.byref(nsv), ctypes.byref(d))
        self.kernel_type, self.gamma, self.rho = kt.value, g.value, rho.value
        self.labels = (lab[0], lab[1])
        self.nsv, self.d = nsv.value, d.value

    def close(self):
        if self.h:
            self._lib.tsvm_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_libsvm(path, labels, x):
    """Dense (labels (n,), features (n, d)) -> LIBSVM text (non-zero features, "%.9g")."""
    lab = np.ascontiguousarray(labels, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float32)
    n, d = xx.shape
    if lab.shape != (n,):
        raise ValueError("labels must have one entry per row")
    st = load().tsvm_write_libsvm(str(path).encode(), _ptr(lab), _ptr(xx), n, d)
    if st != 0:
        raise TsvmError(f"tsvm_write_libsvm({path}) failed with status {st}")


def read_libsvm(path, scale256=False):
    """LIBSVM text -> (labels float32 (n,), dense features float32 (n, d))."""
    L = load()
    lab, inst, n, d = _P(), _P(), _I64(), _I64()
    st = L.tsvm_read_libsvm(path.encode(), int(scale256), ctypes.byref(lab), ctypes.byref(inst),
                            ctypes.byref(n), ctypes.byref(d))
    if st != 0:
        raise TsvmError(f"tsvm_read_libsvm({path}) failed with status {st}")
    nn, dd = n.value, d.value
    labels = np.ctypeslib.as_array(ctypes.cast(lab, ctypes.POINTER(ctypes.c_float)), (max(nn, 1),))[:nn].copy()
    x = np.ctypeslib.as_array(ctypes.cast(inst, ctypes.POINTER(ctypes.c_float)),
                              (max(nn * dd, 1),))[:nn * dd].copy().reshape(nn, dd)
    L.tsvm_free(lab)
    L.tsvm_free(inst)
    return labels, x


def map_labels(labels):
    labels = np.ascontiguousarray(labels, dtype=np.float32).copy()
    pos, neg = _INT(), _INT()
    bad = load().tsvm_map_labels(labels.ctypes.data_as(_P), labels.size, ctypes.byref(pos),
                                 ctypes.byref(neg))
    return labels, pos.value, neg.value, bad


def write_model_rbf(path, alpha, labels, inst, C, gamma, b, pos, neg):
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    labels = np.ascontiguousarray(labels, dtype=np.float32)
    inst = np.ascontiguousarray(inst, dtype=np.float32)
    n, d = inst.shape
    st = load().tsvm_write_model_rbf(path.encode(), alpha.ctypes.data_as(_P), labels.ctypes.data_as(_P),
                                     inst.ctypes.data_as(_P), n, d, C, gamma, b, pos, neg)
    if st != 0:
        raise TsvmError(f"tsvm_write_model_rbf failed ({st})")


def write_model_linear(path, w, b, pos, neg):
    w = np.ascontiguousarray(w, dtype=np.float64)
    st = load().tsvm_write_model_linear(path.encode(), w.ctypes.data_as(_P), w.size, b, pos, neg)
    if st != 0:
        raise TsvmError(f"tsvm_write_model_linear failed ({st})")


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(_P)
    if hasattr(a, "data_ptr"):
        return _P(a.data_ptr())
    return _P(int(a))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def gaussrand_fill(count, nthreads=0):
    """Omega values of the reference's gaussrand() stream (bit-identical), float32."""
    out = np.empty(count, dtype=np.float32)
    st = load().tsvm_gaussrand_fill(_ptr(out), count, nthreads)
    if st != 0:
        raise TsvmError(f"tsvm_gaussrand_fill failed ({st})")
    return out


class Context:
    """A libtsvm context (device, stream, workspace)."""

    def __init__(self, device=0):
        self.lib = load()
        h = _P()
        st = self.lib.tsvm_create(device, ctypes.byref(h))
        if st != 0:
            raise TsvmError(f"tsvm_create({device}) failed with status {st}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.tsvm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != 0:
            msg = self.lib.tsvm_last_error(self.h)
            raise TsvmError(f"{what} failed (status {st}): {msg.decode() if msg else ''}")

    def sync(self):
        self._check(self.lib.tsvm_sync(self.h), "tsvm_sync")

    def kernel_product(self, X1, y1, X2, y2, A, gamma, precision=PREC_FP32):
        X1, y1, X2, y2, A = _f32(X1), _f32(y1), _f32(X2), _f32(y2), _f32(A)
        m, d = X1.shape
        n = X2.shape[0]
        r = A.shape[1]
        W = np.empty((m, r), dtype=np.float32)
        self._check(self.lib.tsvm_kernel_product(self.h, _ptr(X1), _ptr(y1), m, _ptr(X2), _ptr(y2),
                                                 n, d, _ptr(A), r, gamma, precision, _ptr(W)),
                    "tsvm_kernel_product")
        return W

    def lra(self, X, y, k, gamma, precision=PREC_FP32, omega=None):
        """Rank-k factor U (n x k float32) with U U^T ~ diag(y) K diag(y)."""
        X, y = _f32(X), _f32(y)
        n, d = X.shape
        om = None if omega is None else _f32(omega).reshape(n, k)
        U = np.empty((n, k), dtype=np.float32)
        info = LraInfo()
        self._check(self.lib.tsvm_lra(self.h, _ptr(X), _ptr(y), n, d, k, gamma, precision,
                                      _ptr(om), _ptr(U), ctypes.byref(info)), "tsvm_lra")
        return U, info

    def lra_f64(self, X, y, k, gamma, precision=PREC_FP32, omega=None, ldu=None):
        """tsvm_lra_f64 (the reference's LRA types): returns (U (n x ldu float64, the first k
        columns written), l1, info)."""
        X, y = _f32(X), _f32(y)
        n, d = X.shape
        ldu = k if ldu is None else ldu
        om = None if omega is None else _f32(omega).reshape(n, k)
        U = np.zeros((n, ldu))
        l1 = _D()
        info = LraInfo()
        self._check(self.lib.tsvm_lra_f64(self.h, _ptr(X), d, _ptr(y), n, d, k, gamma, precision,
                                          _ptr(om), _ptr(U), ldu, ctypes.byref(l1),
                                          ctypes.byref(info)), "tsvm_lra_f64")
        return U, l1.value, info

    def mpc(self, Z, a, C, k=None, prm=None):
        """tsvm_mpc (the reference's mpc types): Z (n x ldz float64, the first k columns
        used), a (n float64), prm an IpmParams or None; returns (X, Xi, info)."""
        Z = np.ascontiguousarray(Z, dtype=np.float64)
        a = np.ascontiguousarray(a, dtype=np.float64)
        n, ldz = Z.shape
        k = ldz if k is None else k
        X, Xi = np.empty(n), np.empty(n)
        info = IpmInfo()
        self._check(self.lib.tsvm_mpc(self.h, _ptr(Z), ldz, _ptr(a), C, _ptr(X), _ptr(Xi), n, k,
                                      ctypes.byref(prm) if prm is not None else None,
                                      ctypes.byref(info)), "tsvm_mpc")
        return X, Xi, info

    def lra_dev(self, X, y, n, d, k, gamma, precision, omega, U):
        info = LraInfo()
        self._check(self.lib.tsvm_lra_dev(self.h, _ptr(X), _ptr(y), n, d, k, gamma, precision,
                                          _ptr(omega), _ptr(U), ctypes.byref(info)), "tsvm_lra_dev")
        return info

    def ipm_dev(self, Z, a, n, k, C, alpha, xi, s=None, maxit=100, stop=STOP_REF, tol=1e-5,
                precision=PREC_FP32, refine=-1, init=0):
        """Dual IPM on device Z (n x k FP32) and labels a; alpha, xi (and s) device FP64."""
        prm = ipm_params(C, 2, maxit, stop, tol, precision, refine, 0, init)
        info = IpmInfo()
        self._check(self.lib.tsvm_ipm_dev(self.h, _ptr(Z), _ptr(a), n, k, ctypes.byref(prm),
                                          _ptr(alpha), _ptr(xi), _ptr(s) if s is not None else None,
                                          ctypes.byref(info)), "tsvm_ipm_dev")
        return info

    def ipm(self, Z, a, C, kernel_type=2, maxit=100, stop=STOP_REF, tol=1e-5,
            precision=PREC_FP32, refine=-1, verbose=0, init=0):
        """Dual IPM on Q = Z Z^T; returns (alpha, xi, s, info)."""
        Z, a = _f32(Z), _f32(a)
        n, k = Z.shape
        al, xi, s = np.empty(n), np.empty(n), np.empty(n)
        prm = ipm_params(C, kernel_type, maxit, stop, tol, precision, refine, verbose, init)
        info = IpmInfo()
        self._check(self.lib.tsvm_ipm(self.h, _ptr(Z), _ptr(a), n, k, ctypes.byref(prm), _ptr(al),
                                      _ptr(xi), _ptr(s), ctypes.byref(info)), "tsvm_ipm")
        return al, xi, s, info

    def train(self, X, y, k=50, C=1.0, gamma=0.0, kernel_type=2, precision=PREC_FP32,
              stop=STOP_REF, tol=1e-5, maxit=100, refine=-1, verbose=0, omega=None, init=0,
              analysis=0, progress=0):
        """Train on X (n x d) with labels +-1; returns (alpha, w (linear) or None, info)."""
        X, y = _f32(X), _f32(y)
        n, d = X.shape
        om = None if omega is None else _f32(omega).reshape(n, k)
        al = np.empty(n)
        w = np.empty(d)
        prm = train_params(k, C, gamma, kernel_type, precision, stop, tol, maxit, refine, verbose, init,
                           analysis, progress)
        info = TrainInfo()
        self._check(self.lib.tsvm_train(self.h, _ptr(X), _ptr(y), n, d, ctypes.byref(prm), _ptr(om),
                                        _ptr(al), _ptr(w), ctypes.byref(info)), "tsvm_train")
        return al, (w if kernel_type == 0 else None), info

    def train_dev(self, X, y, n, d, prm, omega, alpha):
        info = TrainInfo()
        self._check(self.lib.tsvm_train_dev(self.h, _ptr(X), _ptr(y), n, d, ctypes.byref(prm),
                                            _ptr(omega), _ptr(alpha), ctypes.byref(info)),
                    "tsvm_train_dev")
        return info

    def decision_function(self, X, y, alpha, gamma, b, T):
        X, y, T = _f32(X), _f32(y), _f32(T)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        n, d = X.shape
        m = T.shape[0]
        f = np.empty(m)
        self._check(self.lib.tsvm_decision_function(self.h, _ptr(X), _ptr(y), _ptr(alpha), n, d, gamma,
                                                    b, _ptr(T), m, _ptr(f)), "tsvm_decision_function")
        return f

    def set_progress(self, on=True):
        """Progress lines on stderr during long calls on this context."""
        self.lib.tsvm_set_progress(self.h, int(bool(on)))

    def grad_dev(self, X, y, n, d, alpha, b, k, log_gamma, log_scales, omega, Xt=None, yt=None,
                 nt=0, want_grad=True, precision=PREC_FP32):
        """(D, L, g_dual, g_loss, info): the trained dual through the factor and the held-out
        mean smoothed hinge, and their gradients in (log gamma, log_scales); device X, y,
        alpha (FP64), omega, Xt, yt.  precision: the forward products of the factor."""
        sc = np.ascontiguousarray(log_scales, dtype=np.float64)
        vals = np.zeros(2)
        gD = np.zeros(d + 1)
        gL = np.zeros(d + 1)
        info = GradInfo()
        self._check(self.lib.tsvm_grad_dev(
            self.h, _ptr(X), _ptr(y), n, d, _ptr(alpha), b, k, log_gamma,
            sc.ctypes.data_as(_P), _ptr(omega), _ptr(Xt) if nt else None,
            _ptr(yt) if nt else None, nt, precision, vals.ctypes.data_as(_P),
            gD.ctypes.data_as(_P) if want_grad else None,
            gL.ctypes.data_as(_P) if want_grad else None, ctypes.byref(info)), "tsvm_grad_dev")
        return vals[0], vals[1], gD, gL, info

    def fnorm_res_dev(self, X, y, n, d, gamma, U, k):
        """(|K_Y|_F, |K_Y - U U^T|_F) for device X (n x d), y and U (n x k FP32)."""
        a, b = _D(), _D()
        self._check(self.lib.tsvm_fnorm_res_dev(self.h, _ptr(X), _ptr(y), n, d, gamma, _ptr(U), k,
                                                ctypes.byref(a), ctypes.byref(b)),
                    "tsvm_fnorm_res_dev")
        return a.value, b.value

    def predict(self, model, T):
        """(decision values, predicted labels) of the rows T under a Model."""
        T = _f32(T)
        nt, d = T.shape
        f = np.empty(nt)
        pred = np.empty(nt, dtype=np.float32)
        self._check(self.lib.tsvm_predict(self.h, model.h, _ptr(T), nt, d, _ptr(f), _ptr(pred)),
                    "tsvm_predict")
        return f, pred

    def weighted_syrk_dev(self, Z, n, k, w, precision, M, budget=0.0):
        """M (device, k x k FP64) = I + Z^T diag(w) Z for device Z (n x k FP32), w (FP64)."""
        info = SyrkInfo()
        self._check(self.lib.tsvm_weighted_syrk_dev(self.h, _ptr(Z), n, k, _ptr(w), precision,
                                                    budget, _ptr(M), ctypes.byref(info)),
                    "tsvm_weighted_syrk_dev")
        return info

    def kernel_product_noalloc_dev(self, X1, y1, m, X2, y2, n, d, A, r, gamma, precision, W):
        """tsvm_kernel_product_noalloc_dev: the same product with no device allocation."""
        self._check(self.lib.tsvm_kernel_product_noalloc_dev(self.h, _ptr(X1), _ptr(y1), m,
                                                             _ptr(X2), _ptr(y2), n, d, _ptr(A),
                                                             r, gamma, precision, _ptr(W)),
                    "tsvm_kernel_product_noalloc_dev")

    def kernel_product_dev(self, X1, y1, m, X2, y2, n, d, A, r, gamma, precision, W):
        self._check(self.lib.tsvm_kernel_product_dev(self.h, _ptr(X1), _ptr(y1), m, _ptr(X2),
                                                     _ptr(y2), n, d, _ptr(A), r, gamma,
                                                     precision, _ptr(W)),
                    "tsvm_kernel_product_dev")
