"""Test-only CPU restatement of the closure network (qgn_fortran_amd/csrc/cnn.hip).

This module is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's CPU baseline may import it.  The product path never calls it.

The reference has no network: its closure hook is the QG-Leith term
(ARK43_QGLeith.f90:62-79), so this restatement follows the network's own
definition.  The network has 4 layers of 3x3 periodic convolutions,
cin -> 32 -> 32 -> 32 -> cout, with ReLU after the first three, MSE loss and
Adam.  Tensors are (n, H, W, C).  Parameters: for each layer W[co][tap][ci],
then b[co], with tap = 3 (dy + 1) + (dx + 1); tap (dy, dx) reads
in[y + dy, x + dx].

`bf16=True` rounds at the points where the GPU stores BF16:
- the input;
- the weights;
- the hidden activations after ReLU;
- dL/dpred;
- each masked input gradient.
Accumulation stays FP64 here and is FP32 on the GPU.

Parity pins: the analytic gradients against central finite differences
(FP64, bf16=False); the convolution against a direct loop over taps; Adam
against its closed form after one step (m = (1-b1) g, v = (1-b2) g^2, step =
lr sign(g) to within eps).
"""
from __future__ import annotations

import numpy as np

HIDDEN = 32
LAYERS = 4


def shapes(cin, cout):
    ci = [cin, HIDDEN, HIDDEN, HIDDEN]
    co = [HIDDEN, HIDDEN, HIDDEN, cout]
    return list(zip(ci, co))


def nparams(cin, cout):
    return sum(o * 9 * i + o for i, o in shapes(cin, cout))


def unpack(params, cin, cout):
    """List of (W (co, 9, ci), b (co,)) from the flat vector."""
    out, off = [], 0
    for i, o in shapes(cin, cout):
        W = params[off:off + o * 9 * i].reshape(o, 9, i)
        off += o * 9 * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def init_params(cin, cout, seed, last_scale=0.1):
    """He-normal weights (std sqrt(2 / (9 ci))), zero biases; the last layer
    scaled by last_scale."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (i, o) in enumerate(shapes(cin, cout)):
        W = rng.standard_normal((o, 9, i)) * np.sqrt(2.0 / (9 * i))
        if l == LAYERS - 1:
            W *= last_scale
        layers.append((W, np.zeros(o)))
    return pack(layers).astype(np.float32)


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    """bfloat16 bit patterns (uint16) of a, ties to even."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def shifted(x, dy, dx):
    """s[y, x] = x[y + dy, x + dx] on the periodic grid; x is (n, H, W, C)."""
    return np.roll(x, shift=(-dy, -dx), axis=(1, 2))


def conv(x, W):
    """out[p, co] = sum_tap sum_ci W[co, tap, ci] x[p + off(tap), ci]."""
    out = 0.0
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out = out + np.einsum("nhwi,oi->nhwo", shifted(x, dy, dx), W[:, tap, :])
    return out


def conv_t(d, W):
    """Adjoint of conv: dx[p, ci] = sum_tap sum_co W[co, tap, ci] d[p - off(tap), co]."""
    out = 0.0
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out = out + np.einsum("nhwo,oi->nhwi", shifted(d, -dy, -dx), W[:, tap, :])
    return out


def wgrad(x, d):
    """dW[co, tap, ci] = sum_p d[p, co] x[p + off(tap), ci]."""
    co, ci = d.shape[-1], x.shape[-1]
    dW = np.zeros((co, 9, ci))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dW[:, tap, :] = np.einsum("nhwo,nhwi->oi", d, shifted(x, dy, dx))
    return dW


def forward(params, x, cin, cout, bf16_points=False):
    """Returns (pred, activations [input, a1, a2, a3])."""
    q = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    layers = [(q(W), np.asarray(b, dtype=np.float64)) for W, b in unpack(np.asarray(params), cin, cout)]
    a = q(x)
    acts = [a]
    for l, (W, b) in enumerate(layers):
        z = conv(a, W) + b
        if l < LAYERS - 1:
            a = q(np.maximum(z, 0.0))
            acts.append(a)
        else:
            return z, acts


def loss_and_grads(params, x, target, cin, cout, bf16_points=False):
    """MSE (mean over all n H W cout values) and its gradient vector."""
    q = bf16 if bf16_points else (lambda a: np.asarray(a, dtype=np.float64))
    layers = [(q(W), b) for W, b in unpack(np.asarray(params), cin, cout)]
    pred, acts = forward(params, x, cin, cout, bf16_points)
    diff = pred - target
    N = diff.size
    loss = float(np.sum(diff * diff) / N)
    d = q(2.0 * diff / N)
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        grads[l] = (wgrad(acts[l], d), d.sum(axis=(0, 1, 2)))
        if l > 0:
            d = q(conv_t(d, layers[l][0]) * (acts[l] > 0))
    return loss, pack(grads), pred


def adam(params, grads, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step; returns new (params, m, v)."""
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
