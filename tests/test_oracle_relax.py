"""Pins of the oracle's relaxation path (oracle/relax_oracle.c): the
damping-only RHS, the Bogacki-Shampine RK23 relax step (llg.rs:278-436), the
energy breakdown (energy.rs:80-283), the torque metrics and the relax
controller (relax.rs:334-808).

What pins what:
  * llg_rhs_relax: the vector identity m x (m x B) = m (m.B) - B for |m| = 1,
    and its relation to the full llg_rhs (already pinned by closed forms).
  * RK23 step: the macrospin closed form of damping-only motion, third-order
    convergence of the solution and of the embedded error estimate, FSAL
    equivalence, the controller formula.
  * energy: closed forms for uniform states, the discrete energy gradient
    against the field builder for every term (the reference's
    energy-gradient consistency test, tests/validation.rs:223-318, tightened
    with central differences), demag energy of a uniform film
    (tests/validation.rs:618-644), mask independence.
  * relax controller: stop reasons, stage stops, strides, gates and the
    max_err tightening on a macrospin whose end state is known.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

GAMMA = O.GAMMA_E


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def rand_m(n, seed):
    return unit(np.random.default_rng(seed).standard_normal((n, 3)))


# ---------------------------------------------------------------------------
# llg_rhs_relax
# ---------------------------------------------------------------------------
def test_llg_rhs_relax_identities():
    m = rand_m(50, 1)
    b = np.random.default_rng(2).standard_normal((50, 3))
    alpha = 0.37
    k = O.llg_rhs_relax(m, b, GAMMA, alpha)
    mb = (m * b).sum(axis=1, keepdims=True)
    expect = -GAMMA * alpha / (1 + alpha ** 2) * (m * mb - b)
    np.testing.assert_allclose(k, expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    # full RHS = precession part + damping part
    full = O.llg_rhs(m, b, GAMMA, alpha)
    prec = O.llg_rhs(m, b, GAMMA, 0.0) / (1 + alpha ** 2)
    np.testing.assert_allclose(k, full - prec, rtol=0, atol=1e-12 * np.abs(full).max())
    assert np.abs((k * m).sum(axis=1)).max() <= 1e-12 * np.abs(k).max()


# ---------------------------------------------------------------------------
# RK23 relax step on a macrospin: damping-only motion in a fixed field
# ---------------------------------------------------------------------------
B0 = 0.5
ALPHA = 0.5
GRID1 = O.Grid(1, 1, 5e-9, 5e-9, 5e-9)
MAT1 = O.Material(ms=8e5)


def macrospin_exact(theta0, phi0, t):
    """tan(theta/2) = tan(theta0/2) exp(-alpha gamma B t / (1 + alpha^2)), phi fixed."""
    th = 2 * math.atan(math.tan(theta0 / 2) * math.exp(-ALPHA * GAMMA * B0 * t / (1 + ALPHA ** 2)))
    return np.array([math.sin(th) * math.cos(phi0), math.sin(th) * math.sin(phi0), math.cos(th)])


def m0_macro(theta0=1.0, phi0=0.3):
    return np.array([[math.sin(theta0) * math.cos(phi0), math.sin(theta0) * math.sin(phi0),
                      math.cos(theta0)]])


def run_fixed(dt, t_end):
    ctx, rk = O.Context(GRID1), O.RK23(GRID1)
    m = m0_macro().copy()
    p = O.Params(alpha=ALPHA, dt=dt, b_ext=(0.0, 0.0, B0))
    t = 0.0
    nsteps = int(round(t_end / dt))
    for _ in range(nsteps):
        eps, acc, dtu = rk.step(ctx, m, p, MAT1, O.EXCH_ANIS, 1e-30, 0.8, dt, dt)
        assert acc and dtu == dt
        t += dtu
    return m[0], t


def test_rk23_macrospin_closed_form_adaptive():
    ctx, rk = O.Context(GRID1), O.RK23(GRID1)
    m = m0_macro().copy()
    p = O.Params(alpha=ALPHA, dt=1e-13, b_ext=(0.0, 0.0, B0))
    t, t_end = 0.0, 100e-12
    while t < t_end:
        p.dt = min(p.dt, t_end - t)
        eps, acc, dtu = rk.step(ctx, m, p, MAT1, O.EXCH_ANIS, 1e-10, 0.8, 1e-18, 1e-11)
        if acc:
            t += dtu
    exact = macrospin_exact(1.0, 0.3, t)
    assert np.abs(m[0] - exact).max() < 1e-7
    # no precession in relax mode: the azimuth is unchanged
    assert abs(math.atan2(m[0][1], m[0][0]) - 0.3) < 1e-12


def test_rk23_third_order_convergence():
    t_end = 40e-12
    errs = []
    for dt in (2e-12, 1e-12, 0.5e-12):
        m, t = run_fixed(dt, t_end)
        errs.append(np.abs(m - macrospin_exact(1.0, 0.3, t)).max())
    s1, s2 = math.log2(errs[0] / errs[1]), math.log2(errs[1] / errs[2])
    assert 2.7 < s1 < 3.4 and 2.7 < s2 < 3.4, (errs, s1, s2)


def test_rk23_embedded_error_is_third_order_in_dt():
    eps = []
    for dt in (4e-12, 2e-12, 1e-12):
        ctx, rk = O.Context(GRID1), O.RK23(GRID1)
        m = m0_macro().copy()
        p = O.Params(alpha=ALPHA, dt=dt, b_ext=(0.0, 0.0, B0))
        e, _, _ = rk.step(ctx, m, p, MAT1, O.EXCH_ANIS, 1.0, 0.8, 1e-18, 1.0)
        eps.append(e)
    s1, s2 = math.log2(eps[0] / eps[1]), math.log2(eps[1] / eps[2])
    assert 2.7 < s1 < 3.3 and 2.7 < s2 < 3.3, (eps, s1, s2)


def film_case(nx=12, ny=8, dmi=1.2e-3, demag=True, seed=5):
    g = O.Grid(nx, ny, 4e-9, 4e-9, 3e-9)
    mat = O.Material(ms=8e5, a_ex=1.3e-11, k_u=3e5, easy_axis=(0.0, 0.0, 1.0), dmi=dmi, demag=demag)
    return g, mat, rand_m(g.n, seed)


def test_rk23_fsal_matches_fresh_evaluation():
    g, mat, m0 = film_case()
    p1 = O.Params(alpha=0.3, dt=2e-14, b_ext=(0.01, 0.0, 0.02))
    p2 = O.Params(alpha=0.3, dt=2e-14, b_ext=(0.01, 0.0, 0.02))
    ca, ra = O.Context(g), O.RK23(g)
    cb, rb = O.Context(g), O.RK23(g)
    ma, mb = m0.copy(), m0.copy()
    for step in range(3):
        b0 = ca.field_builds()
        ra.step(ca, ma, p1, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-11)
        assert ca.field_builds() - b0 == (4 if step == 0 else 3)   # FSAL after the first
        assert ra.last_b_eff() is not None
        rb.invalidate()
        rb.step(cb, mb, p2, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-11)
        assert np.array_equal(ma, mb) and p1.dt == p2.dt
    # the cached field is the field of the accepted state
    np.testing.assert_array_equal(ra.last_b_eff(), ca.build_h_eff(ma, p1, mat, O.FULL))
    # a change of b_ext disables FSAL
    p1.b_ext = (0.0, 0.0, 0.0)
    b0 = ca.field_builds()
    ra.step(ca, ma, p1, mat, O.FULL, 1e-3, 0.8, 1e-18, 1e-11)
    assert ca.field_builds() - b0 == 4


def test_rk23_controller_reject_and_floor():
    g, mat, m0 = film_case(demag=False)
    ctx, rk = O.Context(g), O.RK23(g)
    m = m0.copy()
    p = O.Params(alpha=0.3, dt=1e-12)
    eps, acc, dtu = rk.step(ctx, m, p, mat, O.EXCH_ANIS_DMI, 1e-12, 0.8, 1e-18, 1e-11)
    assert not acc and np.array_equal(m, m0) and dtu == 1e-12
    assert rk.last_b_eff() is None
    assert p.dt == pytest.approx(1e-12 * 0.8 * (1e-12 / eps) ** (1 / 3), rel=1e-14)
    # at dt_min the step is accepted whatever the error
    p.dt = 1e-15
    eps, acc, _ = rk.step(ctx, m, p, mat, O.EXCH_ANIS_DMI, 1e-30, 0.8, 1e-15, 1e-11)
    assert acc and eps > 1e-30 and p.dt == 1e-15
    # eps = 0 (uniform state along a field with no torque) doubles dt, clamped
    g1 = O.Grid(1, 1, 5e-9, 5e-9, 5e-9)
    c1, r1 = O.Context(g1), O.RK23(g1)
    mz = np.array([[0.0, 0.0, 1.0]])
    p = O.Params(alpha=0.3, dt=3e-12, b_ext=(0.0, 0.0, 1.0))
    eps, acc, _ = r1.step(c1, mz, p, O.Material(ms=8e5), O.EXCH_ANIS, 1e-5, 0.8, 1e-18, 5e-12)
    assert eps == 0.0 and acc and p.dt == 5e-12


# ---------------------------------------------------------------------------
# energy
# ---------------------------------------------------------------------------
def test_energy_closed_forms_uniform():
    g = O.Grid(10, 7, 4e-9, 5e-9, 3e-9)
    v = 4e-9 * 5e-9 * 3e-9
    u = unit([0.2, -0.3, 0.9])
    d = unit([0.6, 0.3, -0.2])
    mat = O.Material(ms=7e5, a_ex=1.1e-11, k_u=4e5, easy_axis=tuple(u), dmi=None, demag=False)
    m = np.tile(d, (g.n, 1))
    bx = (0.03, -0.02, 0.05)
    e = O.compute_energy(O.Context(g), m, mat, bx)
    assert e["exchange"] == 0.0 and e["dmi"] == 0.0 and e["demag"] == 0.0
    assert e["anisotropy"] == pytest.approx(4e5 * (1 - d.dot(u) ** 2) * v * g.n, rel=1e-13)
    assert e["zeeman"] == pytest.approx(-7e5 * d.dot(bx) * v * g.n, rel=1e-13)


def test_energy_exchange_closed_form_spiral():
    """m rotating by dtheta per cell along x: exchange = A dtheta_eff^2/dx^2 V per pair,
    with |m_{i+1} - m_i|^2 = 2 - 2 cos(dtheta)."""
    g = O.Grid(9, 4, 3e-9, 3e-9, 2e-9)
    v = 3e-9 * 3e-9 * 2e-9
    dth = 0.3
    i = np.tile(np.arange(9), 4)
    m = np.stack([np.cos(dth * i), np.zeros(g.n), np.sin(dth * i)], axis=1)
    mat = O.Material(ms=8e5, a_ex=1.5e-11)
    e = O.compute_energy(O.Context(g), m, mat, (0, 0, 0))
    pairs = 8 * 4
    assert e["exchange"] == pytest.approx(1.5e-11 * (2 - 2 * math.cos(dth)) / 9e-18 * v * pairs, rel=1e-12)


def test_energy_gradient_consistency_all_terms():
    """dE = -Ms V B_eff . dm at an interior cell, for every term at once
    (central differences; all terms are at most quadratic in m, so the
    residual is third order in the perturbation)."""
    g = O.Grid(16, 12, 4e-9, 4e-9, 3e-9)
    mat = O.Material(ms=8e5, a_ex=1.3e-11, k_u=2e5, easy_axis=tuple(unit([0.1, 0.2, 1.0])),
                     dmi=2e-3, demag=True)
    p = O.Params(b_ext=(0.02, -0.01, 0.03))
    ctx = O.Context(g)
    m = rand_m(g.n, 11)
    b = ctx.build_h_eff(m, p, mat, O.FULL)
    v = 4e-9 * 4e-9 * 3e-9
    for (i, j) in ((7, 5), (3, 8), (12, 2)):
        idx = j * 16 + i
        m0 = m[idx]
        a = np.array([0.37, -0.24, 0.91])
        d = unit(a - a.dot(m0) * m0)
        eps = 1e-4
        mp, mm = m.copy(), m.copy()
        mp[idx] = unit(m0 + eps * d)
        mm[idx] = unit(m0 - eps * d)
        de = O.compute_total_energy(ctx, mp, mat, p.b_ext) - O.compute_total_energy(ctx, mm, mat, p.b_ext)
        pred = -mat.ms * v * b[idx].dot(mp[idx] - mm[idx])
        assert abs(de - pred) <= 1e-6 * abs(pred), (i, j, de, pred)


def test_dmi_energy_matches_field_without_mask():
    """E_dmi = -1/2 Ms V sum m . B_dmi with B_dmi from the field builder."""
    g, mat, m = film_case(nx=10, ny=9, dmi=2.5e-3, demag=False)
    ctx = O.Context(g)
    p = O.Params()
    bd = ctx.build_h_eff(m, p, mat, O.EXCH_ANIS_DMI) - ctx.build_h_eff(m, p, mat, O.EXCH_ANIS)
    v = 4e-9 * 4e-9 * 3e-9
    e = O.compute_energy(ctx, m, mat, (0, 0, 0))
    assert e["dmi"] == pytest.approx(-0.5 * mat.ms * v * (m * bd).sum(), rel=1e-10)


def test_demag_energy_uniform_film():
    """tests/validation.rs:618-644 (E_demag >= 0) and the thin-film limit
    E ~ mu0 Ms^2 / 2 per volume for m along z (Nzz < 1 at the edges)."""
    g = O.Grid(32, 32, 5e-9, 5e-9, 1e-9)
    m = np.tile([0.0, 0.0, 1.0], (g.n, 1))
    ctx = O.Context(g)
    e = O.compute_energy(ctx, m, O.Material(ms=1.0, demag=True), (0, 0, 0))
    assert e["demag"] >= 0.0
    ms = 8e5
    e = O.compute_energy(ctx, m, O.Material(ms=ms, demag=True), (0, 0, 0))
    ratio = e["demag"] / (0.5 * O.MU0 * ms ** 2 * 5e-9 * 5e-9 * 1e-9 * g.n)
    assert 0.85 < ratio < 1.0
    # in-plane: much smaller
    mx = np.tile([1.0, 0.0, 0.0], (g.n, 1))
    ex = O.compute_energy(ctx, mx, O.Material(ms=ms, demag=True), (0, 0, 0))
    assert 0.0 < ex["demag"] < 0.1 * e["demag"]


def test_energy_mask_semantics():
    g, mat, m = film_case(nx=10, ny=8, dmi=1.5e-3, demag=True)
    ctx = O.Context(g)
    bx = (0.01, 0.02, -0.03)
    full = O.compute_energy(ctx, m, mat, bx)
    ones = O.compute_energy(ctx, m, mat, bx, geom=np.ones(g.n, dtype=bool))
    for k in full:
        assert ones[k] == full[k], k
    geom = np.ones((8, 10), dtype=bool)
    geom[2:5, 3:6] = False
    geom = geom.reshape(-1)
    e1 = O.compute_energy(ctx, m, O.Material(**{**mat.__dict__, "demag": False}), bx, geom=geom)
    m2 = m.copy()
    m2[~geom] = rand_m(int((~geom).sum()), 99)
    e2 = O.compute_energy(ctx, m2, O.Material(**{**mat.__dict__, "demag": False}), bx, geom=geom)
    for k in e1:
        assert e1[k] == e2[k], k      # cells outside the mask do not enter
    # exchange with a mask counts only pairs inside: compare with a direct pair sum
    v = 4e-9 * 4e-9 * 3e-9
    gm = geom.reshape(8, 10)
    mm = m.reshape(8, 10, 3)
    dx_pairs = gm[:, 1:] & gm[:, :-1]
    dy_pairs = gm[1:, :] & gm[:-1, :]
    ex = (((mm[:, 1:] - mm[:, :-1]) ** 2).sum(-1) * dx_pairs).sum() / (4e-9) ** 2 + \
         (((mm[1:] - mm[:-1]) ** 2).sum(-1) * dy_pairs).sum() / (4e-9) ** 2
    assert e1["exchange"] == pytest.approx(mat.a_ex * ex * v, rel=1e-12)


def test_torque_metrics():
    m = rand_m(37, 3)
    b = np.random.default_rng(4).standard_normal((37, 3))
    t = np.linalg.norm(np.cross(m, b), axis=1)
    assert O.torque_metric(m, b, O.TQ_MAX) == pytest.approx(t.max(), rel=1e-14)
    assert O.torque_metric(m, b, O.TQ_MEAN) == pytest.approx(t.mean(), rel=1e-13)
    assert O.torque_metric(m, b, O.TQ_RMS) == pytest.approx(math.sqrt((t ** 2).mean()), rel=1e-13)
    assert O.torque_metric(m, b, O.TQ_MEANSQ) == pytest.approx((t ** 2).mean(), rel=1e-13)


# ---------------------------------------------------------------------------
# relax controller on a macrospin (easy axis z, tilted start)
# ---------------------------------------------------------------------------
MAT_A = O.Material(ms=8e5, k_u=5e5, easy_axis=(0.0, 0.0, 1.0))


def relax_macro(**kw):
    ctx, rk = O.Context(GRID1), O.RK23(GRID1)
    m = m0_macro(0.5, 0.2).copy()
    p = O.Params(alpha=0.5, dt=1e-13)
    s = O.relax_settings(**kw)
    rep = O.relax_with_report(ctx, rk, m, p, MAT_A, O.EXCH_ANIS, s)
    return rep, m, p, s


def test_relax_default_reaches_threshold_and_floor():
    rep, m, p, s = relax_macro()
    assert rep["stop_reason"] == O.STOP_TIGHTEN_FLOOR
    assert rep["last_stage_stop"] == O.STAGE_BELOW_THRESHOLD
    assert rep["final_max_err"] == 1e-9 and s.max_err == 1e-9
    assert rep["final_torque"] <= 1e-4 and rep["final_torque_max"] <= 1e-4
    assert rep["final_dt"] == p.dt
    assert rep["torque_field_rebuilds"] == 0        # every check follows an accepted step
    assert m[0][2] > 1 - 1e-8
    # tightening 1e-5 -> 1e-9 by 1/sqrt(2): ceil(log(1e4)/log(sqrt 2)) = 27 stages
    stages = math.ceil(math.log(1e4) / math.log(math.sqrt(2)))
    assert rep["torque_checks"] >= stages + 1


def test_relax_phase2_disabled_and_energy_descent():
    ctx = O.Context(GRID1)
    e0 = O.compute_total_energy(ctx, m0_macro(0.5, 0.2), MAT_A, (0, 0, 0))
    rep, m, p, s = relax_macro(phase2_enabled=0)
    assert rep["stop_reason"] == O.STOP_PHASE2_DISABLED
    assert rep["final_torque"] is None and rep["final_torque_max"] is None
    assert rep["accepted_steps"] % 3 == 0 and rep["accepted_steps"] >= 3
    assert O.compute_total_energy(ctx, m, MAT_A, (0, 0, 0)) < e0


def test_relax_max_accepted_steps():
    rep, m, p, s = relax_macro(max_accepted_steps=5)
    assert rep["stop_reason"] == O.STOP_MAX_ACCEPTED and rep["accepted_steps"] == 5
    rep, m, p, s = relax_macro(max_accepted_steps=5, phase1_enabled=0)
    assert rep["stop_reason"] == O.STOP_MAX_ACCEPTED and rep["accepted_steps"] == 5
    assert rep["last_stage_stop"] == O.STAGE_MAX_ACCEPTED and rep["final_torque"] is not None
    assert rep["torque_field_rebuilds"] == 1       # the first check precedes any step


def test_relax_plateau_mode():
    ctx, rk = O.Context(GRID1), O.RK23(GRID1)
    m = m0_macro(0.5, 0.2).copy()
    p = O.Params(alpha=0.5, dt=1e-13)
    s = O.mumax_like_relax_settings()
    rep = O.relax_with_report(ctx, rk, m, p, MAT_A, O.EXCH_ANIS, s)
    assert rep["stop_reason"] == O.STOP_TIGHTEN_FLOOR
    assert rep["last_stage_stop"] == O.STAGE_PLATEAU
    assert m[0][2] > 1 - 1e-6


def test_relax_torque_gate():
    base, _, _, _ = relax_macro()
    rep, m, p, s = relax_macro(final_torque_gate=1e-300, gate_max_extra_accepted_steps=4)
    assert rep["stop_reason"] == O.STOP_GATE_NOT_SATISFIED
    assert rep["accepted_steps"] == base["accepted_steps"] + 4
    rep, m, p, s = relax_macro(final_torque_gate=1.0, gate_max_extra_accepted_steps=4)
    assert rep["stop_reason"] == O.STOP_TIGHTEN_FLOOR
    assert rep["accepted_steps"] == base["accepted_steps"]
