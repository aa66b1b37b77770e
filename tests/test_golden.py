"""Regression fixtures (tests/golden/, written by tests/golden/make_fixtures.py
with the CPU oracle): the oracle must keep reproducing them, and the GPU path
must match them to the parity bar (FP64, 1e-12 of max |B|)."""
import os

import numpy as np
import pytest

from conftest import rel_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


GRID = (32, 16, 4e-9, 4e-9, 3e-9)
MAT = dict(ms=8e5, a_ex=1.3e-11, k_u=2e5, easy_axis=(0.0, 0.0, 1.0), dmi=1.5e-3, demag=True)
PARAMS = dict(alpha=0.02, dt=1e-13, b_ext=(-24.6e-3, 4.3e-3, 0.0))


def test_oracle_reproduces_fixtures():
    from oracle import oracle as O
    f = _load("field_32x16.npz")
    ctx = O.Context(O.Grid(*GRID))
    b = ctx.build_h_eff(f["m"], O.Params(**PARAMS), O.Material(**MAT), O.FULL)
    assert rel_err(b, f["b"]) <= 1e-13
    r = _load("rk45_step_32x16.npz")
    m = r["m0"].copy()
    p = O.Params(**PARAMS)
    eps, acc, dtu = ctx.rk45_step(m, p, O.Material(**MAT), O.FULL, 1e-5, 0.8, 1e-18, 1e-12)
    np.testing.assert_allclose([eps, float(acc), dtu, p.dt], r["out"], rtol=1e-10)
    assert np.abs(m - r["m1"]).max() <= 1e-13
    g = _load("mg_default_16x16.npz")
    bm = np.zeros_like(g["m"])
    mg = O.MG(O.Grid(16, 16, 5e-9, 5e-9, 3e-9), O.mg_config(warm_start=0))
    mg.add_field(g["m"], bm, 8e5)
    assert rel_err(bm, g["b"]) <= 1e-13
    assert mg.stats()["cycles"] == int(g["stats"][0])


@pytest.mark.gpu
def test_gpu_matches_fixtures(gpu_lib):
    L = gpu_lib
    f = _load("field_32x16.npz")
    ctx = L.Context(L.Grid2D(*GRID), precision=L.F64)
    b = ctx.build_h_eff(f["m"], L.LLGParams(**PARAMS), L.Material(**MAT))
    assert rel_err(b, f["b"]) <= 1e-12
    r = _load("rk45_step_32x16.npz")
    m = np.ascontiguousarray(r["m0"].copy())
    p = L.LLGParams(**PARAMS)
    eps, acc, dtu = ctx.step_llg_rk45_recompute_field_adaptive(m, p, L.Material(**MAT), 1e-5, 0.8, 1e-18, 1e-12)
    np.testing.assert_allclose([eps, float(acc), dtu, p.dt], r["out"], rtol=1e-9)
    assert np.abs(m - r["m1"]).max() <= 1e-12
    g = _load("mg_default_16x16.npz")
    c2 = L.Context(L.Grid2D(16, 16, 5e-9, 5e-9, 3e-9), precision=L.F64)
    c2.set_mg_config(L.mg_config(warm_start=0))
    bm = c2.compute_demag_field_poisson_mg(g["m"], L.Material(ms=8e5, demag=True))
    assert rel_err(bm, g["b"]) <= 1e-12
    assert c2.mg_last_stats()["cycles"] == int(g["stats"][0])
