"""ctypes binding of oracle/libqgn_cpu.so, the FP64 C++/OpenMP restatement of
GetRHS and the ARK43 steppers with MKL DFTI (oracle/qgn_cpu.cpp).

TEST INFRASTRUCTURE ONLY: tests/ check it against the numpy oracle, and
bench.py times it as the CPU baseline.  Nothing in the product imports it.
Arrays are the numpy oracle's: spectra complex128 (nz, ny, nx//2+1).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libqgn_cpu.so")
_lib = None


def available():
    """(True, '') when the library is built and MKL loads, else (False, why)."""
    if not os.path.exists(LIB):
        return False, f"{LIB} is not built"
    return True, ""


def _load():
    global _lib
    if _lib is None:
        lib = C.CDLL(LIB)
        D = C.POINTER(C.c_double)
        lib.qcpu_create.restype = C.c_void_p
        lib.qcpu_create.argtypes = [C.c_int] * 3 + [C.c_double] * 10 + [C.c_int] + [D] * 5
        lib.qcpu_destroy.argtypes = [C.c_void_p]
        lib.qcpu_rhs.argtypes = [C.c_void_p, D, D]
        lib.qcpu_ark43_step.argtypes = [C.c_void_p, D, D, D, C.c_int, C.POINTER(C.c_long)]
        lib.qcpu_last_error.restype = C.c_char_p
        lib.qcpu_set_threads.argtypes = [C.c_int]
        lib.qcpu_max_threads.restype = C.c_int
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Model:
    """One model (p: oracle Params-like object with nx..is_explicit, dt_max)."""

    def __init__(self, p, H, Modes, EVals, uBar, qyBar, threads=None):
        lib = _load()
        if threads:
            lib.qcpu_set_threads(int(threads))
        self.threads = lib.qcpu_max_threads()
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (H, Modes, EVals, uBar, qyBar)]
        dt_max = p.dt_max if p.dt_max is not None else np.inf
        self._h = lib.qcpu_create(p.nx, p.ny, p.nz, p.Lx, p.Ly, p.Htot, p.beta, p.A2, p.A8, p.r0, p.C_d, p.TOL,
                                  dt_max, int(p.is_explicit), *[_d(a) for a in self._keep])
        if not self._h:
            raise RuntimeError("qgn_cpu: " + lib.qcpu_last_error().decode())
        self.p = p
        self.rejects = 0

    def get_rhs(self, q_hat):
        q = np.ascontiguousarray(q_hat, dtype=np.complex128)
        out = np.empty_like(q)
        _load().qcpu_rhs(self._h, _d(q.view(np.float64)), _d(out.view(np.float64)))
        return out

    def ark43_step(self, q_hat, t, dt, leith=True):
        """One accepted step; returns (q_hat, t, dt)."""
        q = np.array(q_hat, dtype=np.complex128, order="C", copy=True)
        tt, dd, rj = C.c_double(t), C.c_double(dt), C.c_long(0)
        rc = _load().qcpu_ark43_step(self._h, _d(q.view(np.float64)), C.byref(tt), C.byref(dd), int(leith),
                                     C.byref(rj))
        if rc:
            raise RuntimeError(f"qgn_cpu: ARK43 step failed ({rc})")
        self.rejects += rj.value
        return q, tt.value, dd.value

    def close(self):
        if self._h:
            _load().qcpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
