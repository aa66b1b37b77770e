"""Pins of the U-Net oracle (oracle/unet_oracle.py) that need no GPU.

* bfloat16 rounding: exact values and nearest-even ties.
* The float64 gradient against central finite differences of the loss, for
  parameters of every layer (the backward formulas, the skip and pooling
  paths and the loss region): a dropped term, a wrong tap flip or a swapped
  operand fails here.
* Convolution against a direct per-pixel sum, pooling and upsampling
  against their definitions.
"""
import numpy as np
import pytest

from oracle import unet_oracle as U


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2.0 ** -7, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 3.0e-39])
    r = U.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 + 2.0 ** -7
    assert r[2] == 1.0                       # tie -> even mantissa
    assert r[3] == 1.0 + 2.0 ** -6           # tie -> even mantissa
    assert r[4] == -2.5 and r[5] == 0.0
    assert np.array_equal(U.bf16_from_bits(U.bf16_bits(v)), r)


def test_conv_pool_up_definitions():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 6, 8, 3))
    w = rng.standard_normal((9, 4, 3))
    b = rng.standard_normal(4)
    y = U.conv(x, w, b)
    for (bb, yy, xx) in [(0, 0, 0), (1, 5, 7), (0, 3, 4), (1, 0, 6)]:
        ref = b.copy()
        for tap in range(9):
            dy, dx = tap // 3 - 1, tap % 3 - 1
            if 0 <= yy + dy < 6 and 0 <= xx + dx < 8:
                ref += w[tap] @ x[bb, yy + dy, xx + dx]
        np.testing.assert_allclose(y[bb, yy, xx], ref, rtol=1e-13, atol=1e-13)
    p = U.pool(x)
    np.testing.assert_allclose(p[1, 2, 3], x[1, 4:6, 6:8].mean(axis=(0, 1)), rtol=1e-14)
    u = U.up(p)
    assert np.array_equal(u[1, 5, 7], p[1, 2, 3])
    # adjoints: <pool x, g> = <x, pool_back g>, <up x, g> = <x, up_back g>
    g = rng.standard_normal(p.shape)
    assert abs((p * g).sum() - (x * U.pool_back(g)).sum()) < 1e-12
    g2 = rng.standard_normal(x.shape)
    assert abs((U.up(p) * g2).sum() - (p * U.up_back(g2)).sum()) < 1e-12


def test_float64_gradient_matches_finite_differences():
    s = U.Shape(batch=2, height=8, width=16, out_channels=3, base=16, valid_width=13, valid_height=7)
    rng = np.random.default_rng(1)
    p = U.init_params(s, 2)
    x = np.zeros((2, 8, 16, 16))
    x[..., :4] = rng.standard_normal((2, 8, 16, 4))
    t = np.zeros((2, 8, 16, s.c_pad))
    t[..., :3] = rng.standard_normal((2, 8, 16, 3))
    loss, g, _ = U.loss_and_grad(s, p, x, t, bf16=False)
    assert g.shape == p.shape and loss > 0
    o = 0
    checked = 0
    for li, (ci, co, _) in enumerate(U.layers(s)):
        nw = 9 * ci * co
        cand = [o + int(k) for k in rng.integers(0, nw, 6)] + [o + nw + int(k) for k in rng.integers(0, co, 2)]
        if li == 0:     # the first layer's live input channels
            cand = [o + (tap * co + c) * ci + i for tap, c, i in [(0, 1, 2), (4, 3, 0), (8, 15, 3), (5, 7, 1)]]
        if li == len(U.layers(s)) - 1:
            cand = [o + (tap * co + c) * ci + i for tap, c, i in [(4, 0, 5), (2, 2, 11), (7, 1, 0)]] + [o + nw + 1]
        for k in cand:
            h = 1e-6 * max(1.0, abs(p[k]))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (U.loss_and_grad(s, pp, x, t, bf16=False)[0] - U.loss_and_grad(s, pm, x, t, bf16=False)[0]) / (2 * h)
            assert abs(fd - g[k]) <= 1e-6 * max(abs(fd), 1e-3 * abs(g).max()), (li, k, fd, g[k])
            checked += 1
        o += nw + co
    assert o == U.param_count(s) and checked > 100


def test_padded_channels_get_no_gradient():
    s = U.Shape(batch=1, height=8, width=8, out_channels=5, base=16)
    p = U.init_params(s, 3)
    rng = np.random.default_rng(4)
    x = np.zeros((1, 8, 8, 16))
    x[..., :4] = rng.standard_normal((1, 8, 8, 4))
    t = rng.standard_normal((1, 8, 8, s.c_pad))
    _, g, _ = U.loss_and_grad(s, p, x, t, bf16=True)
    ci, co, _ = U.layers(s)[0]
    w0 = g[:9 * ci * co].reshape(9, co, ci)
    assert np.all(w0[:, :, 4:] == 0.0)        # zero input channels
    head = U.unpack(s, g)[-1]
    assert np.all(head[0][:, 5:, :] == 0.0) and np.all(head[1][5:] == 0.0)   # outside the loss
