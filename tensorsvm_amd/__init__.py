"""tensorsvm_amd: MI355X (gfx950) TensorSVM training hot path.

The product is the C-ABI shared library ``libtsvm.so`` (HIP kernels + C++ host code,
declared in ``include/tsvm.h``) and the ``tensorsvm-train`` command line built beside
it.  This module is a thin ctypes binding of that C-ABI for Python callers; it never
computes anything itself and raises if the library is missing.
"""
from . import capi  # noqa: F401
from .capi import (PREC_BF16, PREC_FP16, PREC_FP16R2, PREC_FP16X2, PREC_FP32, PREC_FP64, PREC_FP8,
                   Context,
                   Model, TsvmError, gaussrand_fill, library_path, load)

__all__ = ["capi", "Context", "Model", "TsvmError", "gaussrand_fill", "load", "library_path",
           "PREC_FP32", "PREC_BF16", "PREC_FP16X2", "PREC_FP16",
           "PREC_FP16R2", "PREC_FP64", "PREC_FP8"]
