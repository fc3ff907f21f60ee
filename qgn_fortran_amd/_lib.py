"""ctypes binding of libqgn.so (declarations in include/qgn.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libqgn.so")

QGN_OK, QGN_EINVAL, QGN_EHIP, QGN_ENOMEM, QGN_ENONFINITE, QGN_EREJECT = 0, -1, -2, -3, -4, -5
SCHEMES = {"AB3": 0, "RK4": 1, "BSRK32": 2, "ARK43": 3, "ARK43_QGLeith": 4}
PRECISIONS = {"fp64": 0, "mixed": 1}
KERNELS = {"all": -2, "row_jac": 0, "col_inv": 1, "col_rhs": 2, "project": 3, "stage": 4, "update": 5, "errnorm": 6}

# Every symbol include/qgn.h declares.
EXPORTS = [
    "qgn_create", "qgn_destroy", "qgn_initialize", "qgn_set_state", "qgn_get_state", "qgn_time_step",
    "qgn_rhs", "qgn_get_psi", "qgn_fields", "qgn_modal_fields", "qgn_spectra", "qgn_leith", "qgn_get_stats", "qgn_get_stream",
    "qgn_prof_enable", "qgn_prof_read", "qgn_dft_create", "qgn_dft_forward", "qgn_dft_backward",
    "qgn_dft_forward_device", "qgn_dft_backward_device", "qgn_dft_free", "qgn_vertical_modes",
    "qgn_last_error", "qgn_version", "qgn_create_batch", "qgn_batch_set_time", "qgn_batch_step",
    "qgn_batch_get_time", "qgn_ensemble_create", "qgn_ensemble_step", "qgn_ensemble_destroy",
    "qgn_cnn_create", "qgn_cnn_destroy", "qgn_cnn_nparams", "qgn_cnn_get_params", "qgn_cnn_set_params",
    "qgn_cnn_forward", "qgn_cnn_backward", "qgn_cnn_adam", "qgn_cnn_train_step", "qgn_cnn_last_error",
    "qgn_set_closure", "qgn_get_tuning", "qgn_set_tuning", "qgn_dft_compute_noalloc",
    "qgn_prof_read_all", "qgn_dfti_create_descriptor", "qgn_dfti_set_value_int", "qgn_dfti_set_value_ints",
    "qgn_dfti_set_value_double", "qgn_dfti_commit_descriptor", "qgn_dfti_compute_forward",
    "qgn_dfti_compute_backward", "qgn_dfti_free_descriptor",
]


class QgnParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Htot", C.c_double),
        ("beta", C.c_double), ("A2", C.c_double), ("A8", C.c_double),
        ("r0", C.c_double), ("C_d", C.c_double), ("TOL", C.c_double), ("dt_max", C.c_double),
        ("is_explicit", C.c_int), ("scheme", C.c_int), ("precision", C.c_int), ("max_rejects", C.c_int),
    ]


class QgnStats(C.Structure):
    _fields_ = [
        ("steps", C.c_longlong), ("attempts", C.c_longlong), ("rejects", C.c_longlong),
        ("rhs_evals", C.c_longlong), ("last_err", C.c_double), ("err0", C.c_double),
    ]


class QgnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qgn error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libqgn.so from the package directory (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m qgn_fortran_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P, D, I = C.c_void_p, C.POINTER(C.c_double), C.c_int
    sig = {
        "qgn_create": (I, [C.POINTER(P), C.POINTER(QgnParams), D, D, D, D, D]),
        "qgn_destroy": (I, [P]),
        "qgn_initialize": (I, [P, D, D]),
        "qgn_set_state": (I, [P, D]),
        "qgn_get_state": (I, [P, D]),
        "qgn_time_step": (I, [P, D, D]),
        "qgn_rhs": (I, [P, D, D]),
        "qgn_get_psi": (I, [P, D, D]),
        "qgn_fields": (I, [P, D, D]),
        "qgn_modal_fields": (I, [P, D, D, D]),
        "qgn_spectra": (I, [P, I, D, D]),
        "qgn_leith": (I, [P, D]),
        "qgn_get_stats": (I, [P, C.POINTER(QgnStats)]),
        "qgn_get_stream": (I, [P, C.POINTER(P)]),
        "qgn_prof_enable": (I, [P, I]),
        "qgn_prof_read": (I, [P, C.POINTER(I), D]),
        "qgn_dft_create": (I, [C.POINTER(P), I, I, I, I]),
        "qgn_dft_forward": (I, [P, D, D]),
        "qgn_dft_backward": (I, [P, D, D]),
        "qgn_dft_forward_device": (I, [P, P, P, P]),
        "qgn_dft_backward_device": (I, [P, P, P, P]),
        "qgn_dft_compute_noalloc": (I, [I, I, I, I, I, P, P, P]),
        "qgn_prof_read_all": (I, [P, C.POINTER(C.c_int), D]),
        "qgn_dfti_create_descriptor": (I, [C.POINTER(P), I, I, I, C.POINTER(C.c_longlong)]),
        "qgn_dfti_set_value_int": (I, [P, I, C.c_longlong]),
        "qgn_dfti_set_value_ints": (I, [P, I, C.POINTER(C.c_longlong), I]),
        "qgn_dfti_set_value_double": (I, [P, I, C.c_double]),
        "qgn_dfti_commit_descriptor": (I, [P]),
        "qgn_dfti_compute_forward": (I, [P, D, D]),
        "qgn_dfti_compute_backward": (I, [P, D, D]),
        "qgn_dfti_free_descriptor": (I, [C.POINTER(P)]),
        "qgn_dft_free": (I, [P]),
        "qgn_vertical_modes": (I, [I, C.c_double, D, D, D, D, D, D, D, D]),
        "qgn_create_batch": (I, [C.POINTER(P), C.POINTER(QgnParams), I, D, D, D, D, D, D]),
        "qgn_batch_set_time": (I, [P, D, D]),
        "qgn_batch_step": (I, [P, I]),
        "qgn_batch_get_time": (I, [P, D, D, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(I)]),
        "qgn_ensemble_create": (I, [C.POINTER(P), C.POINTER(P), I]),
        "qgn_ensemble_step": (I, [P, I]),
        "qgn_ensemble_destroy": (I, [P]),
        "qgn_cnn_create": (I, [C.POINTER(P), I, I, C.POINTER(C.c_float)]),
        "qgn_cnn_destroy": (I, [P]),
        "qgn_cnn_nparams": (I, [P, C.POINTER(I)]),
        "qgn_cnn_get_params": (I, [P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qgn_cnn_set_params": (I, [P, C.POINTER(C.c_float)]),
        "qgn_cnn_forward": (I, [P, P, I, I, I, P]),
        "qgn_cnn_backward": (I, [P, P, D]),
        "qgn_cnn_adam": (I, [P, C.c_float, C.c_float, C.c_float, C.c_float]),
        "qgn_cnn_train_step": (I, [P, P, P, I, I, I, P, C.c_float, D]),
        "qgn_cnn_last_error": (C.c_char_p, []),
        "qgn_set_closure": (I, [P, P, D, D]),
        "qgn_get_tuning": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_float)]),
        "qgn_set_tuning": (I, [P, I, I, I]),
        "qgn_last_error": (C.c_char_p, []),
        "qgn_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = _lib.qgn_last_error().decode() if _lib is not None else ""
        raise QgnError(code, msg)
    return code


def check_cnn(code):
    if code != 0:
        msg = _lib.qgn_cnn_last_error().decode() if _lib is not None else ""
        raise QgnError(code, msg)
    return code


def fptr(a):
    """float* view of a C-contiguous float32 numpy array."""
    if a is None:
        return None
    if a.dtype != "float32" or not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous float32")
    return a.ctypes.data_as(C.POINTER(C.c_float))


def dptr(a):
    """double* view of a C-contiguous float64/complex128 numpy array."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.POINTER(C.c_double))
