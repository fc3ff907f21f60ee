"""Pins of the closure-network oracle (oracle/cnn_oracle.py) on the CPU:
finite differences for every parameter block, a direct loop for the
periodic convolution, the adjoint identity of conv/conv_t, the bf16 rounding
against known bit patterns and Adam's first step in closed form."""
import numpy as np
import pytest

from oracle import cnn_oracle as co


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 3))
    W = rng.standard_normal((4, 9, 3))
    out = co.conv(x, W)
    n, H, Wd, _ = x.shape
    ref = np.zeros((n, H, Wd, 4))
    for b in range(n):
        for y in range(H):
            for xx in range(Wd):
                for tap in range(9):
                    dy, dx = tap // 3 - 1, tap % 3 - 1
                    ref[b, y, xx] += W[:, tap, :] @ x[b, (y + dy) % H, (xx + dx) % Wd]
    assert np.allclose(out, ref, rtol=1e-13, atol=1e-13)


def test_conv_t_is_the_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 7, 5, 3))
    d = rng.standard_normal((2, 7, 5, 4))
    W = rng.standard_normal((4, 9, 3))
    lhs = np.sum(co.conv(x, W) * d)
    rhs = np.sum(x * co.conv_t(d, W))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    # wgrad is the gradient of <conv(x, W), d> in W
    g = co.wgrad(x, d)
    E = rng.standard_normal(W.shape)
    assert abs(np.sum(co.conv(x, E) * d) - np.sum(g * E)) < 1e-10


@pytest.mark.parametrize("cin,cout", [(16, 8), (32, 16)])
def test_gradients_match_finite_differences(cin, cout):
    rng = np.random.default_rng(2)
    n, H, W = 2, 6, 5
    p = co.init_params(cin, cout, seed=3, last_scale=1.0).astype(np.float64)
    p[np.arange(p.size) % 7 == 0] += 0.05      # non-zero biases too
    x = rng.standard_normal((n, H, W, cin))
    t = rng.standard_normal((n, H, W, cout))
    loss, g, _ = co.loss_and_grads(p, x, t, cin, cout)
    # sample entries from every layer's W and b
    offs, idx = 0, []
    for i, o in co.shapes(cin, cout):
        nw = o * 9 * i
        idx += list(rng.integers(offs, offs + nw, 6)) + list(rng.integers(offs + nw, offs + nw + o, 2))
        offs += nw + o
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        lp, _, _ = co.loss_and_grads(p + e, x, t, cin, cout)
        lm, _, _ = co.loss_and_grads(p - e, x, t, cin, cout)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_bf16_rounding_bit_patterns():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-39, 65504.0], dtype=np.float32)
    bits = co.bf16_bits(vals)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up to 1+2^-7
    assert list(bits[:4]) == [0x3F80, 0x3F80, 0x3F81, 0xC020]
    assert np.all(co.bf16(vals) == (bits.astype(np.uint32) << 16).view(np.float32))


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    new, m, v = co.adam(p, g, np.zeros(50), np.zeros(50), 1, lr=1e-3)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
    assert np.allclose(new, p - 1e-3 * np.sign(g), atol=1e-9)
