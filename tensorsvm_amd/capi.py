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
    _bind_optional(L)
    _LIB = L
    return L


def _bind_optional(L):
    """Bind entry points added beside the kernel product (same library)."""
    for name, args, res in _EXTRA:
        fn = getattr(L, name, None)
        if fn is not None:
            fn.argtypes = args
            if res is not None:
                fn.restype = res


_EXTRA = []


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

    def kernel_product_dev(self, X1, y1, m, X2, y2, n, d, A, r, gamma, precision, W):
        self._check(self.lib.tsvm_kernel_product_dev(self.h, _ptr(X1), _ptr(y1), m, _ptr(X2),
                                                     _ptr(y2), n, d, _ptr(A), r, gamma,
                                                     precision, _ptr(W)),
                    "tsvm_kernel_product_dev")
