"""CPU restatement of the learned multigrid initial guess (the U-Net of
llg-sim_amd/csrc/unet.hip): forward pass, loss and the gradient of every
parameter, in numpy float64.

TEST INFRASTRUCTURE ONLY: the checker of the HIP network.  The reference has
no learned component (its multigrid solve starts from zero or from the
previous solve, demag_poisson_mg.rs:1947-2057); what this restates is the
network the library defines, so that the device's forward and backward
kernels are checked against an independent implementation.

With bf16=True the values are rounded to bfloat16 (nearest even) where the
device stores bf16: the input, the weights, every activation, each pooled
value, and the back-propagated error before it enters a product.  With
bf16=False everything is float64: that mode is pinned by central finite
differences (tests/test_oracle_unet.py), which fixes the backward formulas.

Parameter vector (the library's layout): for each convolution in the order
enc 0..3 (two each), dec 2..0 (two each), head: weights [tap][c_out][c_in]
(tap = (dy+1)*3 + (dx+1)) then bias [c_out].
"""
from dataclasses import dataclass
from typing import Dict, List, Tuple

import numpy as np

LEVELS = 4


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Nearest-even rounding to bfloat16 (via float32), returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a: np.ndarray) -> np.ndarray:
    """uint16 bit patterns of bf16_round(a)."""
    return (bf16_round(a).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def bf16_from_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@dataclass
class Shape:
    batch: int
    height: int
    width: int
    out_channels: int
    base: int = 32
    in_channels: int = 4
    valid_width: int = 0
    valid_height: int = 0

    @property
    def c_pad(self) -> int:
        return (self.out_channels + 15) // 16 * 16

    def c(self, l: int) -> int:
        return self.base << l


def layers(s: Shape) -> List[Tuple[int, int, bool]]:
    """(c_in, c_out, relu) of every convolution, in parameter order."""
    out = []
    for l in range(LEVELS):
        out.append((16 if l == 0 else s.c(l - 1), s.c(l), True))
        out.append((s.c(l), s.c(l), True))
    for l in range(LEVELS - 2, -1, -1):
        out.append((s.c(l + 1) + s.c(l), s.c(l), True))
        out.append((s.c(l), s.c(l), True))
    out.append((s.c(0), s.c_pad, False))
    return out


def unpack(s: Shape, p: np.ndarray) -> List[Tuple[np.ndarray, np.ndarray]]:
    res, o = [], 0
    for ci, co, _ in layers(s):
        w = p[o:o + 9 * co * ci].reshape(9, co, ci)
        o += 9 * co * ci
        b = p[o:o + co]
        o += co
        res.append((w, b))
    assert o == p.size, (o, p.size)
    return res


def param_count(s: Shape) -> int:
    return sum(9 * ci * co + co for ci, co, _ in layers(s))


def _shift(x: np.ndarray, dy: int, dx: int) -> np.ndarray:
    """out[b, y, x] = x[b, y+dy, x+dx], zero outside."""
    out = np.zeros_like(x)
    H, W = x.shape[1], x.shape[2]
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray) -> np.ndarray:
    """y[b,y,x,o] = sum_{tap,i} x[b, y+dy, x+dx, i] w[tap, o, i] + bias[o]."""
    y = np.zeros(x.shape[:3] + (w.shape[1],))
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        y += _shift(x, dy, dx) @ w[tap].T
    return y + b


def conv_back(x: np.ndarray, w: np.ndarray, dz: np.ndarray, need_dx: bool = True):
    """Gradients of conv: (dx, dw, db) for output error dz."""
    dw = np.zeros_like(w)
    dx = np.zeros_like(x) if need_dx else None
    flat = dz.reshape(-1, dz.shape[-1])
    for tap in range(9):
        dy_, dx_ = tap // 3 - 1, tap % 3 - 1
        dw[tap] = flat.T @ _shift(x, dy_, dx_).reshape(-1, x.shape[-1])
        if need_dx:
            dx += _shift(dz, -dy_, -dx_) @ w[tap]
    return dx, dw, flat.sum(0)


def pool(x: np.ndarray) -> np.ndarray:
    return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])


def pool_back(g: np.ndarray) -> np.ndarray:
    return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)


def up(x: np.ndarray) -> np.ndarray:
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(g: np.ndarray) -> np.ndarray:
    return g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]


def forward(s: Shape, p: np.ndarray, x: np.ndarray, bf16: bool = True) -> Tuple[np.ndarray, Dict]:
    """x: [B, H, W, 16] float64 (already bf16 values when bf16).  Returns the
    FP32-valued output [B, H, W, c_pad] and the saved activations."""
    R = bf16_round if bf16 else (lambda a: a)
    prm = unpack(s, p)
    wq = [(R(w), b) for w, b in prm]
    act: Dict = {"x": x}
    k = 0
    h = x
    skips = []
    for l in range(LEVELS):
        if l > 0:
            h = R(pool(skips[-1]))
            act[f"pool{l}"] = h
        act[f"ein{l}"] = h
        h = R(np.maximum(conv(h, *wq[k]), 0.0)); k += 1
        act[f"ea{l}"] = h
        h = R(np.maximum(conv(h, *wq[k]), 0.0)); k += 1
        act[f"eb{l}"] = h
        skips.append(h)
    below = skips[-1]
    for l in range(LEVELS - 2, -1, -1):
        cat = np.concatenate([up(below), skips[l]], axis=-1)
        act[f"cat{l}"] = cat
        h = R(np.maximum(conv(cat, *wq[k]), 0.0)); k += 1
        act[f"da{l}"] = h
        h = R(np.maximum(conv(h, *wq[k]), 0.0)); k += 1
        act[f"db{l}"] = h
        below = h
    out = conv(below, *wq[k])
    act["wq"] = wq
    return out, act


def loss_and_grad(s: Shape, p: np.ndarray, x: np.ndarray, t: np.ndarray, bf16: bool = True):
    """Mean squared error over channels < out_channels and the valid region,
    and its gradient w.r.t. the parameter vector (same layout as p)."""
    R = bf16_round if bf16 else (lambda a: a)
    out, a = forward(s, p, x, bf16)
    wq = a["wq"]
    vw = min(s.valid_width or s.width, s.width)
    vh = min(s.valid_height or s.height, s.height)
    sel = np.zeros(out.shape, dtype=bool)
    sel[:, :vh, :vw, :s.out_channels] = True
    count = s.batch * vw * vh * s.out_channels
    d = np.where(sel, out - t, 0.0)
    loss = float((d * d).sum() / count)
    g_out = 2.0 * d / count
    grads = [None] * len(wq)
    k = len(wq) - 1

    def step(k, xin, dz, need_dx=True):
        dx, dw, db = conv_back(xin, wq[k][0], dz, need_dx)
        grads[k] = (dw, db)
        return dx

    g = step(k, a["db0"], R(g_out))            # head
    k -= 1
    g_skip = {}
    for l in range(0, LEVELS - 1):
        g = step(k, a[f"da{l}"], R(g * (a[f"db{l}"] > 0))); k -= 1
        g = step(k, a[f"cat{l}"], R(g * (a[f"da{l}"] > 0))); k -= 1
        cu = s.c(l + 1)
        g_skip[l] = g[..., cu:]
        g = up_back(g[..., :cu])
    g_eb = {LEVELS - 1: g}
    for l in range(LEVELS - 1, -1, -1):
        ge = g_eb[l] if l == LEVELS - 1 else g_skip[l] + g_eb.get(l, 0.0)
        kk = 2 * l
        g = step(kk + 1, a[f"ea{l}"], R(ge * (a[f"eb{l}"] > 0)))
        if l == 0:
            step(kk, a["ein0"], R(g * (a["ea0"] > 0)), need_dx=False)
        else:
            g = step(kk, a[f"pool{l}"], R(g * (a[f"ea{l}"] > 0)))
            g_eb[l - 1] = pool_back(g)
    flat = np.concatenate([np.concatenate([dw.reshape(-1), db]) for dw, db in grads])
    return loss, flat, out


def init_params(s: Shape, seed: int) -> np.ndarray:
    """Random parameters for tests (not the library's initialiser)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (ci, co, _) in enumerate(layers(s)):
        fan = 9 * (s.in_channels if i == 0 else ci)
        w = rng.standard_normal((9, co, ci)) * np.sqrt(2.0 / fan)
        if i == 0:
            w[:, :, s.in_channels:] = 0.0
        parts += [w.reshape(-1), 0.1 * rng.standard_normal(co)]
    return np.concatenate(parts)
