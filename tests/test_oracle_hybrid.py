"""Pins of the oracle's hybrid MG demag (oracle/hybrid_oracle.c):
MG + local delta-kernel correction (demag_poisson_mg.rs:270-700, 2277-2530,
3159-3597).

* Gaussian screening: normalised weights of exp(-x^2 / 2 sigma^2) on
  [-ceil(3 sigma), ceil(3 sigma)]; the screened rhs equals a separable
  convolution with edge clamping (scipy.ndimage), boundary rows zeroed.
* The reference's own test demag_hybrid_reduces_random_error_vs_pure_mg
  (tests/test_demag_methods.rs:197-267): on its 32 x 32 random state the
  hybrid must reach rel-RMSE < 0.12 against FFT and < 0.35 of the pure-MG
  error; the author's recorded run is "~0.32 -> ~0.04".  That test sets the
  radius but not LLG_DEMAG_MG_HYBRID_ENABLE, and its recorded figure is the
  unscreened hybrid (sigma = 0): with sigma = 0 and radius 2 the oracle gives
  0.047 against 0.328 for pure MG.  With the current default screening
  (sigma = 1.5 cells) a radius of 2 cannot hold the screened part (0.29); a
  radius of 5 can (0.027).
* The delta kernel's uniform-bias fix: the diagonal stencils sum to zero.
* A uniform state sees no correction from the interior stencil (bias fix).
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def test_gaussian_weights():
    for sigma in (0.5, 1.0, 1.5, 2.7):
        w = O.gaussian_kernel_1d(sigma)
        r = math.ceil(3 * sigma)
        x = np.arange(-r, r + 1)
        want = np.exp(-x * x / (2 * sigma * sigma))
        np.testing.assert_allclose(w, want / want.sum(), rtol=1e-15, atol=0)
    assert O.gaussian_kernel_1d(0.0).tolist() == [1.0]


def test_screen_rhs_is_separable_gaussian_with_clamped_edges():
    g = O.Grid(12, 10, 4e-9, 4e-9, 3e-9)
    mg = O.MG(g, O.mg_config(pad_xy=3, n_vac_z=2))
    rhs = np.random.default_rng(4).standard_normal((mg.pz, mg.py, mg.px))
    mg.set_finest("rhs", rhs)
    sigma = 1.3
    mg.screen_rhs(sigma)
    got = mg.finest("rhs")
    w = O.gaussian_kernel_1d(sigma)
    # x pass on interior rows (others zero), clamped at the row ends; then y pass
    tmp = ndimage.convolve1d(rhs, w[::-1], axis=2, mode="nearest")
    tmp[:, :, 0] = 0.0
    tmp[:, :, -1] = 0.0
    tmp[[0, -1], :, :] = 0.0
    tmp[:, [0, -1], :] = 0.0
    want = ndimage.convolve1d(tmp, w[::-1], axis=1, mode="nearest")
    want[:, :, 0] = 0.0
    want[:, :, -1] = 0.0
    want[[0, -1], :, :] = 0.0
    want[:, [0, -1], :] = 0.0
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 * np.abs(rhs).max())


def _xorshift_unit_vectors(n, seed):
    """init_random_unit_vectors of tests/test_demag_methods.rs:144-180."""
    mask = (1 << 64) - 1
    x = seed
    out = np.zeros((n, 3))

    def nxt():
        nonlocal x
        x ^= (x << 13) & mask
        x ^= x >> 7
        x ^= (x << 17) & mask
        return x

    for c in range(n):
        v = [2.0 * (nxt() / float(mask)) - 1.0 for _ in range(3)]
        nrm = max(math.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2), 1e-30)
        out[c] = [v[0] / nrm, v[1] / nrm, v[2] / nrm]
    return out


def _rel_rmse(ref, test):
    return math.sqrt(((ref - test) ** 2).sum() / (ref ** 2).sum())


def test_reference_hybrid_reduces_random_error_vs_pure_mg():
    """tests/test_demag_methods.rs:197-267 (with the hybrid switched on, as
    the recorded run was): 32 x 32, 5 x 5 x 1 nm, random unit m."""
    g = O.Grid(32, 32, 5e-9, 5e-9, 1e-9)
    m = _xorshift_unit_vectors(32 * 32, 12345)
    mat = O.Material(ms=8e5, demag=True, demag_method=1)
    bf = np.zeros_like(m)
    O.Demag(g).add_field(m, bf, 8e5)
    cfg = O.mg_config(warm_start=0, pad_xy=2, n_vac_z=16, v_cycles=4)
    pure = O.Context(g)
    pure.set_mg_config(cfg)
    b0 = np.zeros_like(m)
    pure.add_demag(m, b0, mat)
    rel0 = _rel_rmse(bf, b0)
    assert 0.2 < rel0 < 0.45, rel0
    for sigma, radius, lo, hi in ((0.0, 2, 0.03, 0.06), (1.5, 5, 0.0, 0.05)):
        hyb = O.Context(g)
        hyb.set_mg_config(cfg)
        hyb.set_hyb_config(O.hyb_config(enabled=1, radius_xy=radius, delta_v_cycles=20, sigma_cells=sigma))
        bh = np.zeros_like(m)
        hyb.add_demag(m, bh, mat)
        relh = _rel_rmse(bf, bh)
        assert lo < relh < hi, (sigma, radius, relh, rel0)
        assert relh < 0.12 and relh < 0.35 * rel0, (relh, rel0)


def test_delta_kernel_bias_fix_and_uniform_state():
    g = O.Grid(24, 20, 4e-9, 4e-9, 2e-9)
    cfg = O.mg_config(warm_start=0, pad_xy=3, n_vac_z=4, v_cycles=3)
    h = O.hyb_config(enabled=1, radius_xy=3, delta_v_cycles=12)
    r = O.hyb_effective_radius(g, h)
    assert r == 3
    dk = O.hyb_delta_kernel(g, cfg, h, r, 8e5)
    for c in (0, 2, 3):
        assert abs(dk[c].sum()) <= 1e-12 * np.abs(dk[c]).max()
    # a uniform state gets no correction where the whole stencil is inside
    m = np.tile([0.3, -0.4, np.sqrt(1 - 0.25)], (g.n, 1))
    b = np.zeros_like(m)
    O.hyb_add_correction(g, dk, m, b, 8e5)
    bb = b.reshape(20, 24, 3)[r:-r, r:-r]
    off = dk[1].sum() * 8e5 * np.array([m[0, 1], m[0, 0], 0.0])   # the xy stencil is not bias-fixed
    assert np.abs(bb - off).max() <= 1e-9 * np.abs(dk).max() * 8e5
    # effective radius clamps to the grid
    assert O.hyb_effective_radius(O.Grid(5, 30, 1e-9, 1e-9, 1e-9), O.hyb_config(enabled=1, radius_xy=12)) == 1


def test_hybrid_disabled_is_pure_mg():
    g = O.Grid(16, 12, 4e-9, 4e-9, 3e-9)
    m = np.random.default_rng(2).standard_normal((g.n, 3))
    mat = O.Material(ms=8e5, demag=True, demag_method=1)
    cfg = O.mg_config(warm_start=0, pad_xy=2, n_vac_z=2, v_cycles=3)
    a, b = O.Context(g), O.Context(g)
    a.set_mg_config(cfg)
    b.set_mg_config(cfg)
    b.set_hyb_config(O.hyb_config(enabled=1, radius_xy=0))     # radius 0: disabled
    ba, bb = np.zeros_like(m), np.zeros_like(m)
    a.add_demag(m, ba, mat)
    b.add_demag(m, bb, mat)
    np.testing.assert_array_equal(ba, bb)
