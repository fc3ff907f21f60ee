"""The learned subgrid closure network over the HIP library (include/qgn.h,
qgn_cnn_*): a 4-layer 3x3 periodic CNN with BF16 MFMA forward/backward
kernels and Adam.  Device tensors are torch CUDA tensors (used only as
device memory): inputs as int16 bit patterns of bfloat16, outputs and
targets float32, all (n, H, W, C) with the channel fastest."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check_cnn, fptr

HIDDEN = 32


def nparams(cin, cout):
    dims = [(cin, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, cout)]
    return sum(o * 9 * i + o for i, o in dims)


def to_bf16_bits(a):
    """bfloat16 bit patterns (int16) of a float array, round to nearest even."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16).view(np.int16)


class CNN:
    def __init__(self, cin, cout, params):
        self.lib = _lib.load()
        self.cin, self.cout = cin, cout
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != nparams(cin, cout):
            raise ValueError(f"expected {nparams(cin, cout)} parameters, got {p.size}")
        self._h = C.c_void_p()
        check_cnn(self.lib.qgn_cnn_create(C.byref(self._h), cin, cout, fptr(p)))
        self.np = p.size

    def params(self):
        out = np.empty(self.np, dtype=np.float32)
        check_cnn(self.lib.qgn_cnn_get_params(self._h, fptr(out), None))
        return out

    def grads(self):
        out = np.empty(self.np, dtype=np.float32)
        check_cnn(self.lib.qgn_cnn_get_params(self._h, None, fptr(out)))
        return out

    def set_params(self, p):
        check_cnn(self.lib.qgn_cnn_set_params(self._h, fptr(np.ascontiguousarray(p, dtype=np.float32))))

    @staticmethod
    def _check(t, name, dtype, shape):
        """A contiguous CUDA tensor of the given dtype and shape: the library
        reads and writes through its raw pointer."""
        import torch
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise ValueError(f"{name} must be a CUDA tensor")
        if t.dtype != dtype:
            raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def _shape(self, x, y=None):
        import torch
        if x.dim() != 4:
            raise ValueError("x must be (n, H, W, cin)")
        n, H, W, _ = x.shape
        self._check(x, "x", torch.int16, (n, H, W, self.cin))
        if y is not None:
            self._check(y, "y", torch.float32, (n, H, W, self.cout))
        return n, H, W

    @staticmethod
    def _ready(t):
        """The library runs on its own stream: let the work torch queued on
        the tensor's current stream (which may still be writing it) finish
        first.  The library synchronises its own stream before returning."""
        import torch
        torch.cuda.current_stream(t.device).synchronize()

    def forward(self, x, y):
        """x: int16 CUDA tensor (n, H, W, cin); y: float32 CUDA tensor (n, H, W, cout)."""
        n, H, W = self._shape(x, y)
        self._last = (n, H, W)
        self._ready(x)
        check_cnn(self.lib.qgn_cnn_forward(self._h, C.c_void_p(x.data_ptr()), n, H, W, C.c_void_p(y.data_ptr())))

    def backward(self, target):
        """Gradients of the mean-squared error of the last forward's output
        against target; that forward's x and y must still be alive and
        unchanged."""
        import torch
        if getattr(self, "_last", None) is None:
            raise ValueError("backward needs a forward first")
        self._check(target, "target", torch.float32, self._last + (self.cout,))
        loss = C.c_double()
        self._ready(target)
        check_cnn(self.lib.qgn_cnn_backward(self._h, C.c_void_p(target.data_ptr()), C.byref(loss)))
        return loss.value

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        check_cnn(self.lib.qgn_cnn_adam(self._h, lr, b1, b2, eps))

    def train_step(self, x, target, y, lr):
        import torch
        n, H, W = self._shape(x, y)
        self._check(target, "target", torch.float32, (n, H, W, self.cout))
        self._last = (n, H, W)
        loss = C.c_double()
        self._ready(x)
        check_cnn(self.lib.qgn_cnn_train_step(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(target.data_ptr()),
                                              n, H, W, C.c_void_p(y.data_ptr()), lr, C.byref(loss)))
        return loss.value

    def close(self):
        """Destroy the network; raises (and keeps it) while a model has it
        attached or an ensemble holds it."""
        if self._h:
            check_cnn(self.lib.qgn_cnn_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
