"""Pins the oracle's Dormand-Prince RK45 stepper (llg.rs:795-1026) on the CPU:

* damped macrospin in a constant field against the closed form
  phi = gamma B t / (1+a^2), tan(theta/2) = tan(theta0/2) exp(-a gamma B t/(1+a^2));
* fixed-step (dt_min = dt_max) global error slope ~5 (5th-order solution) and
  error-estimate slope ~5 (the embedded pair differs at O(dt^5) per step): a
  wrong tableau entry drops one of the orders;
* FSAL: 7 field builds on the first step, 6 per accepted step after it, 7
  again after a reject or a change of B_ext (llg.rs:854-867, 1018-1023);
* reject path leaves m unchanged and shrinks dt;
* reference test macrospin_precession_quarter_turn_rk45_adaptive
  (tests/validation.rs:646-745).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def closed_form(theta0, t, gamma, b0, alpha):
    f = gamma * b0 / (1 + alpha * alpha)
    th = 2 * math.atan(math.tan(theta0 / 2) * math.exp(-alpha * f * t))
    ph = f * t
    return np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])


def run_fixed(dt, t_end, alpha=0.1, b0=0.5, theta0=1.0):
    c = O.Context(O.Grid(1, 1, 1e-9, 1e-9, 1e-9))
    m = np.array([[math.sin(theta0), 0.0, math.cos(theta0)]])
    p = O.Params(alpha=alpha, dt=dt, b_ext=(0.0, 0.0, b0))
    mat = O.Material(ms=8e5)
    n = int(round(t_end / dt))
    eps_max = 0.0
    for _ in range(n):
        eps, acc, _ = c.rk45_step(m, p, mat, O.FULL, 1e-30, 0.8, dt, dt)
        assert acc
        eps_max = max(eps_max, eps)
    return m[0], eps_max


def test_damped_macrospin_matches_closed_form_adaptive():
    c = O.Context(O.Grid(1, 1, 1e-9, 1e-9, 1e-9))
    theta0, b0, alpha = 1.2, 0.2, 0.05
    m = np.array([[math.sin(theta0), 0.0, math.cos(theta0)]])
    p = O.Params(alpha=alpha, dt=1e-13, b_ext=(0.0, 0.0, b0))
    t, t_end = 0.0, 2e-10
    while t < t_end:
        p.dt = min(p.dt, t_end - t)
        _, acc, dtu = c.rk45_step(m, p, O.Material(ms=8e5), O.FULL, 1e-10, 0.8, 1e-20, 1e-11)
        if acc:
            t += dtu
    ref = closed_form(theta0, t, O.GAMMA_E, b0, alpha)
    assert np.abs(m[0] - ref).max() < 1e-7


def test_fixed_step_convergence_order_is_five():
    t_end = 4e-11
    errs, epss = [], []
    dts = [2e-12, 1e-12, 5e-13]
    for dt in dts:
        mm, eps = run_fixed(dt, t_end)
        errs.append(np.abs(mm - closed_form(1.0, t_end, O.GAMMA_E, 0.5, 0.1)).max())
        epss.append(eps)
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    slope_eps = np.polyfit(np.log(dts), np.log(epss), 1)[0]
    assert 4.5 < slope < 5.6, (slope, errs)
    assert 4.5 < slope_eps < 5.6, (slope_eps, epss)


def test_fsal_field_build_counts():
    g = O.Grid(8, 4, 4e-9, 4e-9, 3e-9)
    c = O.Context(g)
    m = np.random.default_rng(0).standard_normal((32, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    mat = O.Material(ms=8e5, a_ex=1.3e-11)
    p = O.Params(alpha=0.1, dt=1e-14, b_ext=(0.0, 0.0, 0.01))
    c.rk45_step(m, p, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-14)
    assert c.field_builds() == 7
    c.rk45_step(m, p, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-14)
    assert c.field_builds() == 13
    p.b_ext = (0.0, 0.0, 0.02)
    c.rk45_step(m, p, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-14)
    assert c.field_builds() == 20
    c.rk45_step(m, p, mat, O.EXCH_ANIS, 1e-3, 0.8, 1e-18, 1e-14)
    assert c.field_builds() == 27


def test_reject_leaves_state_and_shrinks_dt():
    g = O.Grid(8, 4, 4e-9, 4e-9, 3e-9)
    c = O.Context(g)
    m = np.random.default_rng(1).standard_normal((32, 3))
    m /= np.linalg.norm(m, axis=1, keepdims=True)
    m0 = m.copy()
    p = O.Params(alpha=0.1, dt=1e-11)
    eps, acc, dtu = c.rk45_step(m, p, O.Material(ms=8e5, a_ex=1.3e-11), O.FULL, 1e-9, 0.8, 1e-18, 1e-11)
    assert not acc and dtu == 1e-11 and (m == m0).all() and p.dt < 1e-11
    # the next attempt recomputes stage 1 (no FSAL after a reject)
    before = c.field_builds()
    c.rk45_step(m, p, O.Material(ms=8e5, a_ex=1.3e-11), O.FULL, 1e-9, 0.8, 1e-18, 1e-11)
    assert c.field_builds() - before == 7


def test_macrospin_precession_quarter_turn_rk45_adaptive():
    c = O.Context(O.Grid(1, 1, 1.0, 1.0, 1.0))
    m = np.array([[1.0, 0.0, 0.0]])
    b0 = 0.1
    p = O.Params(alpha=0.0, dt=1e-14, b_ext=(0.0, 0.0, b0))
    t_target = (math.pi / 2) / (p.gamma * b0)
    t, attempts = 0.0, 0
    while t < t_target:
        attempts += 1
        assert attempts < 200000
        if p.dt > t_target - t:
            p.dt = t_target - t
        _, acc, dtu = c.rk45_step(m, p, O.Material(ms=8e5), O.FULL, 1e-5, 0.8, 1e-18, 1e-11)
        if acc:
            t += dtu
    v = m[0]
    assert abs(v[2]) < 0.1 and abs(v[1]) > 0.9 and abs(v[0]) < 0.3
