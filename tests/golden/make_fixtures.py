"""Writes the regression fixtures of tests/golden/ with the CPU oracle.

The reference (Rust) cannot be built or run in this environment and ships no
vectors for this path, so these are the oracle's own outputs on small seeded
inputs, frozen so that a change to the oracle or to the GPU path that moves a
result shows up.  What pins the oracle itself is tests/test_oracle_*.py.

    python tests/golden/make_fixtures.py      (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

GRID = (32, 16, 4e-9, 4e-9, 3e-9)
MAT = dict(ms=8e5, a_ex=1.3e-11, k_u=2e5, easy_axis=(0.0, 0.0, 1.0), dmi=1.5e-3, demag=True)
PARAMS = dict(alpha=0.02, dt=1e-13, b_ext=(-24.6e-3, 4.3e-3, 0.0))


def unit_m(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def main():
    g = O.Grid(*GRID)
    m = unit_m(g.n, 2024)
    ctx = O.Context(g)
    b = ctx.build_h_eff(m, O.Params(**PARAMS), O.Material(**MAT), O.FULL)
    np.savez(os.path.join(HERE, "field_32x16.npz"), m=m, b=b)

    m_rk = m.copy()
    p = O.Params(**PARAMS)
    eps, acc, dtu = ctx.rk45_step(m_rk, p, O.Material(**MAT), O.FULL, 1e-5, 0.8, 1e-18, 1e-12)
    np.savez(os.path.join(HERE, "rk45_step_32x16.npz"), m0=m, m1=m_rk,
             out=np.array([eps, float(acc), dtu, p.dt]))

    gm = O.Grid(16, 16, 5e-9, 5e-9, 3e-9)
    mm = unit_m(gm.n, 7)
    mm[[0, 17, 34]] = 0.0
    cfg = O.mg_config(warm_start=0)
    bm = np.zeros_like(mm)
    mg = O.MG(gm, cfg)
    mg.add_field(mm, bm, 8e5)
    st = mg.stats()
    np.savez(os.path.join(HERE, "mg_default_16x16.npz"), m=mm, b=bm,
             stats=np.array([st["cycles"], st["last_residual"], st["tol_target"]]))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
