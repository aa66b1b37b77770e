"""GPU parity of the Dormand-Prince RK45 stepper (llg.rs:795-1026) through the C ABI.

Bars: FP64 context: the accepted state, eps and next dt of single steps within
1e-12 of the oracle; spatially averaged m(t) within 1e-4 of the oracle over a
trajectory (dt sequence and accept/reject decisions must follow the oracle's).
Reference test mirrored: macrospin_precession_quarter_turn_rk45_adaptive
(tests/validation.rs:646-745).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(nx, ny, seed, dmi=None, demag=True, k_u=0.0):
    from oracle import oracle as O
    import llgmi as L
    from initial_states import smoothed_random
    m0 = smoothed_random(nx, ny, seed).reshape(-1, 3).copy()
    kw = dict(ms=8.0e5, a_ex=1.3e-11, k_u=k_u, dmi=dmi, demag=demag)
    return m0, O.Material(**kw), L.Material(**kw)


def test_rk45_single_steps_match_oracle(gpu_lib):
    from oracle import oracle as O
    L = gpu_lib
    nx, ny, cell = 32, 16, (4e-9, 4e-9, 3e-9)
    m0, om, gm = _setup(nx, ny, 5, dmi=1e-3, k_u=2e4)
    oc = O.Context(O.Grid(nx, ny, *cell))
    gc = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    op = O.Params(alpha=0.02, dt=1e-13, b_ext=(-24.6e-3, 4.3e-3, 0.0))
    gp = L.LLGParams(alpha=0.02, dt=1e-13, b_ext=(-24.6e-3, 4.3e-3, 0.0))
    mo, mg = m0.copy(), m0.copy()
    for step in range(6):
        eo, ao, uo = oc.rk45_step(mo, op, om, 2, 1e-6, 0.8, 1e-19, 1e-11)
        eg, ag, ug = gc.step_llg_rk45_recompute_field_masked_adaptive(mg, gp, gm, 2, 1e-6, 0.8, 1e-19, 1e-11)
        assert ao == ag, step
        assert uo == ug
        assert abs(eg - eo) <= 1e-9 * abs(eo) + 1e-30, (step, eg, eo)
        assert abs(gp.dt - op.dt) <= 1e-9 * op.dt
        assert np.abs(mg - mo).max() <= 1e-12, step
        # keep both on the oracle's dt so later steps compare like with like
        gp.dt = op.dt


def test_rk45_reject_path_and_fsal(gpu_lib):
    """A huge first dt is rejected (state unchanged, dt reduced); FSAL is only
    used after an accepted step with the same mask and B_ext."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny, cell = 16, 8, (4e-9, 4e-9, 3e-9)
    m0, om, gm = _setup(nx, ny, 9)
    oc = O.Context(O.Grid(nx, ny, *cell))
    gc = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    op, gp = O.Params(alpha=0.1, dt=5e-12), L.LLGParams(alpha=0.1, dt=5e-12)
    mo, mg = m0.copy(), m0.copy()
    eo, ao, _ = oc.rk45_step(mo, op, om, 2, 1e-7, 0.8, 1e-18, 1e-11)
    eg, ag, _ = gc.step_llg_rk45_recompute_field_masked_adaptive(mg, gp, gm, 2, 1e-7, 0.8, 1e-18, 1e-11)
    assert not ao and not ag
    assert (mg == m0).all()
    assert abs(gp.dt - op.dt) <= 1e-9 * op.dt
    for _ in range(3):
        gp.dt = op.dt
        oc.rk45_step(mo, op, om, 2, 1e-7, 0.8, 1e-18, 1e-11)
        gc.step_llg_rk45_recompute_field_masked_adaptive(mg, gp, gm, 2, 1e-7, 0.8, 1e-18, 1e-11)
    assert np.abs(mg - mo).max() <= 1e-12
    # changing B_ext invalidates FSAL in both
    op.b_ext = gp.b_ext = (0.0, 0.0, 0.05)
    gp.dt = op.dt
    oc.rk45_step(mo, op, om, 2, 1e-7, 0.8, 1e-18, 1e-11)
    gc.step_llg_rk45_recompute_field_masked_adaptive(mg, gp, gm, 2, 1e-7, 0.8, 1e-18, 1e-11)
    assert np.abs(mg - mo).max() <= 1e-12


def _trajectory(stepper, m, params, t_end, dt_out):
    """Driver of bin/uniform_film_field_rk45.rs:316-343: clamp dt to land on
    each output time; returns <m> at every output time."""
    out = []
    t = 0.0
    n_out = int(round(t_end / dt_out))
    for k in range(1, n_out + 1):
        t_target = k * dt_out
        while t + 1e-18 < t_target:
            remaining = t_target - t
            if params.dt > remaining:
                params.dt = remaining
            eps, acc, dtu = stepper(m, params)
            if acc:
                t += dtu
        t = t_target
        out.append(m.reshape(-1, 3).mean(axis=0).copy())
    return np.array(out)


def test_rk45_trajectory_matches_oracle(gpu_lib):
    """Plain film (exchange, Zeeman, FFT demag), 0.1 ns, FP64 on both sides."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny, cell = 64, 32, (4e-9, 4e-9, 3e-9)
    m0, om, gm = _setup(nx, ny, 1)
    bext = (-24.6e-3, 4.3e-3, 0.0)
    oc = O.Context(O.Grid(nx, ny, *cell))
    gc = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    op = O.Params(alpha=0.02, dt=1e-13, b_ext=bext)
    gp = L.LLGParams(alpha=0.02, dt=1e-13, b_ext=bext)
    args = (1e-6, 0.8, 1e-19, 1e-11)
    ro = _trajectory(lambda m, p: oc.rk45_step(m, p, om, 2, *args), m0.copy(), op, 1e-10, 1e-11)
    rg = _trajectory(lambda m, p: gc.step_llg_rk45_recompute_field_masked_adaptive(m, p, gm, 2, *args),
                     m0.copy(), gp, 1e-10, 1e-11)
    assert np.abs(rg - ro).max() <= 1e-4


def test_macrospin_precession_quarter_turn_rk45_adaptive(gpu_lib):
    """tests/validation.rs:646-745 through the C ABI, plus the closed form."""
    L = gpu_lib
    grid = L.Grid2D(1, 1, 1.0, 1.0, 1.0)
    b0 = 0.1
    params = L.LLGParams(alpha=0.0, dt=1e-14, b_ext=(0.0, 0.0, b0))
    mat = L.Material(ms=8.0e5)
    ctx = L.Context(grid, precision=L.F64)
    m = np.array([[1.0, 0.0, 0.0]])
    t_target = (math.pi / 2) / (params.gamma * b0)
    t, attempts = 0.0, 0
    while t < t_target:
        attempts += 1
        assert attempts < 200000
        remaining = t_target - t
        if params.dt > remaining:
            params.dt = remaining
        _, acc, dtu = ctx.step_llg_rk45_recompute_field_adaptive(m, params, mat, 1e-5, 0.8, 1e-18, 1e-11)
        if acc:
            t += dtu
    v = m[0]
    assert abs(v[2]) < 0.1 and abs(v[1]) > 0.9 and abs(v[0]) < 0.3
    # closed form: dm/dt = -gamma m x B with m = x, B = b0 z gives +gamma b0 y,
    # so m = (cos phi, sin phi, 0) with phi = gamma b0 t
    phi = params.gamma * b0 * t
    np.testing.assert_allclose(v, [math.cos(phi), math.sin(phi), 0.0], atol=1e-4)


def test_rk45_fp32_matches_fp64_over_steps(gpu_lib):
    """FP32 context (FP64 error norm + controller) against the FP64 context."""
    L = gpu_lib
    nx, ny, cell = 64, 32, (4e-9, 4e-9, 3e-9)
    m0, _, gm = _setup(nx, ny, 2)
    c64 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    c32 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    bext = (-24.6e-3, 4.3e-3, 0.0)
    p64 = L.LLGParams(alpha=0.02, dt=1e-13, b_ext=bext)
    p32 = L.LLGParams(alpha=0.02, dt=1e-13, b_ext=bext)
    args = (1e-6, 0.8, 1e-19, 1e-11)
    r64 = _trajectory(lambda m, p: c64.step_llg_rk45_recompute_field_adaptive(m, p, gm, *args), m0.copy(), p64, 5e-11, 1e-11)
    r32 = _trajectory(lambda m, p: c32.step_llg_rk45_recompute_field_adaptive(m, p, gm, *args), m0.copy(), p32, 5e-11, 1e-11)
    assert np.abs(r32 - r64).max() <= 1e-4


def _films(seed, k):
    """k small independent films with their own sizes and materials."""
    import llgmi as L
    from initial_states import smoothed_random
    rng = np.random.default_rng(seed)
    films = []
    for i in range(k):
        nx, ny = int(rng.choice([8, 16, 24, 32])), int(rng.choice([8, 12, 16]))
        cell = (float(rng.uniform(2e-9, 5e-9)), float(rng.uniform(2e-9, 5e-9)), 3e-9)
        mat = L.Material(ms=float(rng.uniform(4e5, 1.2e6)), a_ex=float(rng.uniform(5e-12, 2e-11)),
                         k_u=float(rng.uniform(0, 5e5)), dmi=float(rng.uniform(0, 4e-3)), demag=True)
        par = L.LLGParams(alpha=float(rng.uniform(0.01, 0.5)), dt=1e-13, b_ext=(0.0, 0.0, 0.01 * i))
        films.append((L.Grid2D(nx, ny, *cell), mat, par, smoothed_random(nx, ny, 100 + i).reshape(-1, 3)))
    return films


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rk45_batch_equals_films_stepped_alone(gpu_lib, prec):
    """llgmi_dev_rk45_step_batch gives every film exactly (bitwise) what
    llgmi_dev_rk45_step gives it alone: state, eps, accept and both dts."""
    from conftest import DeviceBuffer
    import copy
    L = gpu_lib
    P = L.F64 if prec == "f64" else L.F32
    es = 8 if prec == "f64" else 4
    films = _films(3, 6)
    ctx_b = [L.Context(g, precision=P) for g, _, _, _ in films]
    ctx_s = [L.Context(g, precision=P) for g, _, _, _ in films]
    buf_b = [DeviceBuffer(3 * g.n_cells() * es) for g, _, _, _ in films]
    buf_s = [DeviceBuffer(3 * g.n_cells() * es) for g, _, _, _ in films]
    for (g, _, _, m0), cb, cs, bb, bs in zip(films, ctx_b, ctx_s, buf_b, buf_s):
        cb.dev_upload(m0, bb)
        cs.dev_upload(m0, bs)
    pb = [copy.deepcopy(f[2]) for f in films]
    ps = [copy.deepcopy(f[2]) for f in films]
    pb[1].dt = ps[1].dt = 5e-11          # one film starts with a step that is rejected
    mats = [f[1] for f in films]
    n_acc = 0
    for attempt in range(8):
        res = L.dev_rk45_step_batch(ctx_b, buf_b, pb, mats, 2, 1e-5, 0.8, 1e-18, 1e-11)
        for i in range(len(films)):
            alone = ctx_s[i].dev_rk45_step(buf_s[i], ps[i], mats[i], 2, 1e-5, 0.8, 1e-18, 1e-11)
            assert res[i] == alone, (attempt, i)
            assert pb[i].dt == ps[i].dt
            n_acc += res[i][1]
            if attempt == 0 and i == 1:
                assert not res[i][1]
    assert n_acc > 30
    for cb, cs, bb, bs in zip(ctx_b, ctx_s, buf_b, buf_s):
        assert np.array_equal(cb.dev_download(bb), cs.dev_download(bs))


def test_rk45_batch_films_match_oracle(gpu_lib):
    """Per-film accuracy of the batch against the FP64 oracle (FP64 contexts,
    the single-step bars: state 1e-12, eps and dt 1e-9 relative)."""
    from conftest import DeviceBuffer
    from oracle import oracle as O
    L = gpu_lib
    films = _films(11, 3)
    ctxs = [L.Context(g, precision=L.F64) for g, _, _, _ in films]
    bufs = [DeviceBuffer(3 * g.n_cells() * 8) for g, _, _, _ in films]
    ocs, oms, ops, mos = [], [], [], []
    for (g, mat, par, m0), c, b in zip(films, ctxs, bufs):
        c.dev_upload(m0, b)
        ocs.append(O.Context(O.Grid(g.nx, g.ny, g.dx, g.dy, g.dz)))
        oms.append(O.Material(ms=mat.ms, a_ex=mat.a_ex, k_u=mat.k_u, dmi=mat.dmi, demag=True))
        ops.append(O.Params(alpha=par.alpha, dt=par.dt, b_ext=par.b_ext))
        mos.append(m0.copy())
    pars = [f[2] for f in films]
    for step in range(4):
        res = L.dev_rk45_step_batch(ctxs, bufs, pars, [f[1] for f in films], 2, 1e-6, 0.8, 1e-19, 1e-11)
        for i in range(len(films)):
            eo, ao, uo = ocs[i].rk45_step(mos[i], ops[i], oms[i], 2, 1e-6, 0.8, 1e-19, 1e-11)
            eg, ag, ug = res[i]
            assert ag == ao and ug == uo, (step, i)
            assert abs(eg - eo) <= 1e-9 * abs(eo) + 1e-30
            assert abs(pars[i].dt - ops[i].dt) <= 1e-9 * ops[i].dt
            pars[i].dt = ops[i].dt
            assert np.abs(ctxs[i].dev_download(bufs[i]) - mos[i]).max() <= 1e-12, (step, i)


def test_rk45_batch_rejects_bad_input_without_stepping(gpu_lib):
    from conftest import DeviceBuffer
    L = gpu_lib
    films = _films(5, 3)
    ctxs = [L.Context(g, precision=L.F64) for g, _, _, _ in films]
    bufs = [DeviceBuffer(3 * g.n_cells() * 8) for g, _, _, _ in films]
    for (g, _, _, m0), c, b in zip(films, ctxs, bufs):
        c.dev_upload(m0, b)
    pars = [f[2] for f in films]
    mats = [f[1] for f in films]
    with pytest.raises(L.LlgmiError, match="twice"):
        L.dev_rk45_step_batch([ctxs[0], ctxs[1], ctxs[0]], bufs, pars, mats, 2, 1e-6, 0.8, 1e-19, 1e-11)
    bad = list(mats)
    bad[2] = L.Material(ms=8e5, a_ex=1e-11, demag_method=7)
    with pytest.raises(L.LlgmiError, match="film 2"):
        L.dev_rk45_step_batch(ctxs, bufs, pars, bad, 2, 1e-6, 0.8, 1e-19, 1e-11)
    small = DeviceBuffer(8)
    with pytest.raises(L.LlgmiError, match="film 1"):
        L.dev_rk45_step_batch(ctxs, [bufs[0], small, bufs[2]], pars, mats, 2, 1e-6, 0.8, 1e-19, 1e-11)
    assert [p.dt for p in pars] == [1e-13] * 3
    for (g, _, _, m0), c, b in zip(films, ctxs, bufs):
        assert np.array_equal(c.dev_download(b), m0)
    assert L.dev_rk45_step_batch([], [], [], [], 2, 1e-6, 0.8, 1e-19, 1e-11) == []


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rk45_graph_replay_equals_eager(gpu_lib, prec, monkeypatch):
    """Attempts replayed from captured HIP graphs (the default on the FFT
    path) against the same attempts launched eagerly (LLGMI_RK45_GRAPH=0),
    step by step: state, eps, accept and dt bitwise equal over 12 attempts,
    including a rejected one and a B_ext change (new capture), with the two
    contexts' eager work and synchronisations interleaved between replays."""
    from conftest import DeviceBuffer
    L = gpu_lib
    P, es = (L.F32, 4) if prec == "f32" else (L.F64, 8)
    nx, ny, cell = 48, 32, (3e-9, 3e-9, 2e-9)
    m0, _, mat = _setup(nx, ny, 17, dmi=2e-3, k_u=3e4)
    cg = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    ce = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    bg, be = DeviceBuffer(3 * nx * ny * es), DeviceBuffer(3 * nx * ny * es)
    cg.dev_upload(m0, bg)
    ce.dev_upload(m0, be)
    pg = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    pe = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    for step in range(12):
        if step == 5:
            pg.dt = pe.dt = 4e-11                # rejected
        if step == 8:
            pg.b_ext = pe.b_ext = (0.01, 0.0, 0.02)
        monkeypatch.delenv("LLGMI_RK45_GRAPH", raising=False)
        rg = cg.dev_rk45_step(bg, pg, mat, 2, 1e-5, 0.8, 1e-18, 1e-11)
        monkeypatch.setenv("LLGMI_RK45_GRAPH", "0")
        re_ = ce.dev_rk45_step(be, pe, mat, 2, 1e-5, 0.8, 1e-18, 1e-11)
        assert rg == re_, step
        assert pg.dt == pe.dt
        assert np.array_equal(cg.dev_download(bg), ce.dev_download(be)), step
        if step == 5:
            assert not rg[1]


@pytest.mark.parametrize("path", ["fft_unfused", "fft_fused", "fft_radix3", "no_demag", "mg"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_rk45_stage_epilogue_equals_stage_kernels(gpu_lib, prec, path, monkeypatch):
    """The next stage state formed in the epilogue of the field evaluation
    (the default) against a stage kernel of its own per stage
    (LLGMI_RK45_FUSE=0): state, eps, accept and dt bitwise equal over 10
    attempts (one rejected, one B_ext change) and over a device-controlled run
    of 6 attempts; with the FFT demag (inverse row pass fused with the local
    field or not; a 384 x 8 film whose rows pad to 768 = 3 * 256, the radix-3
    transforms), without demag, and with the multigrid demag."""
    from conftest import DeviceBuffer
    L = gpu_lib
    P, es = (L.F32, 4) if prec == "f32" else (L.F64, 8)
    nx, ny, cell = (384, 8, (3e-9, 3e-9, 2e-9)) if path == "fft_radix3" else (40, 24, (3e-9, 3e-9, 2e-9))
    if path == "fft_radix3":
        monkeypatch.setenv("LLGMI_FFT_PAD", "3")
    m0, _, mat = _setup(nx, ny, 23, dmi=2e-3, k_u=3e4, demag=path != "no_demag")
    if path == "mg":
        mat.demag_method = L.DemagMethod.PoissonMG
    mask = 1 if path == "no_demag" else 2
    monkeypatch.setenv("LLGMI_FFT_FUSED", "1" if path == "fft_fused" else "0")
    ca = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    cb = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    ba, bb = DeviceBuffer(3 * nx * ny * es), DeviceBuffer(3 * nx * ny * es)
    ca.dev_upload(m0, ba)
    cb.dev_upload(m0, bb)
    pa = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    pb = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    for step in range(10):
        if step == 4:
            pa.dt = pb.dt = 4e-11                # rejected
        if step == 7:
            pa.b_ext = pb.b_ext = (0.01, 0.0, 0.02)
        monkeypatch.delenv("LLGMI_RK45_FUSE", raising=False)
        ra = ca.dev_rk45_step(ba, pa, mat, mask, 1e-5, 0.8, 1e-18, 1e-11)
        monkeypatch.setenv("LLGMI_RK45_FUSE", "0")
        rb = cb.dev_rk45_step(bb, pb, mat, mask, 1e-5, 0.8, 1e-18, 1e-11)
        assert ra == rb, step
        assert pa.dt == pb.dt
        assert np.array_equal(ca.dev_download(ba), cb.dev_download(bb)), step
        if step == 4:
            assert not ra[1]
    if path == "mg":
        return   # the device-controlled run is built for the FFT demag
    monkeypatch.delenv("LLGMI_RK45_FUSE", raising=False)
    oa = ca.dev_rk45_run(ba, pa, mat, mask, 1e-5, 0.8, 1e-18, 1e-11, attempts=6)
    monkeypatch.setenv("LLGMI_RK45_FUSE", "0")
    ob = cb.dev_rk45_run(bb, pb, mat, mask, 1e-5, 0.8, 1e-18, 1e-11, attempts=6)
    assert oa == ob
    assert pa.dt == pb.dt
    assert np.array_equal(ca.dev_download(ba), cb.dev_download(bb))


def test_c1_one_nanosecond_trajectory_matches_oracle(gpu_lib):
    """The 128 x 64 film of 4 x 4 x 3 nm cells (Ms 8e5, A 1.3e-11, alpha 0.02,
    B = (-24.6, 4.3, 0) mT, FFT demag padded to 256 x 128), seeded
    box-smoothed random m0, adaptive RK45 with max_err 1e-6 for 1 ns, outputs
    every 10 ps: the FP64 path's spatially averaged m stays within 1e-4 of the
    CPU oracle's (measured 5.4e-10) with the same number of attempts (4252).
    The FP32 path follows the FP64 one within 1e-4 for the first 150 ps; from
    random m0 the film then forms domains, which amplify FP32 rounding into
    different textures (measured 2e-4 at 300 ps, 5e-2 by 1 ns)."""
    from conftest import DeviceBuffer
    from oracle import oracle as O
    L = gpu_lib
    nx, ny, cell = 128, 64, (4e-9, 4e-9, 3e-9)
    m0, om, mat = _setup(nx, ny, 4)
    bext = (-24.6e-3, 4.3e-3, 0.0)

    def run(step, m_get, p, n_out):
        avg, t, n = [], 0.0, 0
        for k in range(1, n_out + 1):
            t_target = k * 1e-11
            while t + 1e-18 < t_target:
                if p.dt > t_target - t:
                    p.dt = t_target - t
                _, acc, dtu = step(p)
                n += 1
                if acc:
                    t += dtu
            t = t_target
            avg.append(m_get().mean(axis=0))
        return np.array(avg), n

    gpu = {}
    for prec, es, n_out in ((L.F64, 8, 100), (L.F32, 4, 15)):
        ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=prec)
        buf = DeviceBuffer(3 * nx * ny * es)
        ctx.dev_upload(m0, buf)
        p = L.LLGParams(alpha=0.02, dt=1e-13, b_ext=bext)
        gpu[prec] = run(lambda q, c=ctx, b=buf: c.dev_rk45_step(b, q, mat, 2, 1e-6, 0.8, 1e-19, 1e-11),
                        lambda c=ctx, b=buf: c.dev_download(b), p, n_out)
    oc = O.Context(O.Grid(nx, ny, *cell))
    mo = m0.copy()
    ro, no = run(lambda q: oc.rk45_step(mo, q, om, 2, 1e-6, 0.8, 1e-19, 1e-11), lambda: mo,
                 O.Params(alpha=0.02, dt=1e-13, b_ext=bext), 100)
    d64 = np.abs(gpu[L.F64][0] - ro).max()
    d32 = np.abs(gpu[L.F32][0] - gpu[L.F64][0][:15]).max()
    print(f"1 ns FP64 vs oracle {d64:.2e} ({gpu[L.F64][1]} / {no} attempts); 150 ps FP32 vs FP64 {d32:.2e}")
    assert d64 <= 1e-4 and gpu[L.F64][1] == no
    assert d32 <= 1e-4


def test_c2_fp32_rk45_against_fp64_over_200_steps(gpu_lib):
    """The 1024 x 1024 workload (2 x 2 x 1 nm cells, Ms 5.8e5, A 1.5e-11,
    Ku 8e5 along z, D 3e-3, alpha 0.3, B = (-24.6, 4.3, 0) mT, FFT demag
    2048^2, +-z Voronoi domains with 10 % noise): 200 accepted FP64 RK45
    steps (max_err 1e-5), then the FP32 path driven to the same time with its
    own controller.  Bars: <m> within 1e-6 (measured 1.0e-9), every cell
    within 1e-4 (4.3e-6), and the FP32 effective field of the final state
    within 1e-5 of FP64's, max |dB| / max |B| (2.1e-7)."""
    from conftest import DeviceBuffer
    from initial_states import voronoi_domains
    L = gpu_lib
    n = 1024
    cell = (2e-9, 2e-9, 1e-9)
    mat = L.Material(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, easy_axis=(0.0, 0.0, 1.0), dmi=3e-3, demag=True)
    bext = (-24.6e-3, 4.3e-3, 0.0)
    m0 = voronoi_domains(n, n, seed=2048, n_points=2048, noise=0.1).reshape(-1, 3)
    c64 = L.Context(L.Grid2D(n, n, *cell), precision=L.F64)
    c32 = L.Context(L.Grid2D(n, n, *cell), precision=L.F32)
    b64, b32 = DeviceBuffer(3 * n * n * 8), DeviceBuffer(3 * n * n * 4)
    c64.dev_upload(m0, b64)
    c32.dev_upload(m0, b32)
    p64 = L.LLGParams(alpha=0.3, dt=1e-14, b_ext=bext)
    t64, acc = 0.0, 0
    while acc < 200:
        _, a, u = c64.dev_rk45_step(b64, p64, mat, 2, 1e-5, 0.8, 1e-18, 1e-12)
        if a:
            acc += 1
            t64 += u
    p32 = L.LLGParams(alpha=0.3, dt=1e-14, b_ext=bext)
    t32 = 0.0
    while t32 + 1e-22 < t64:
        if p32.dt > t64 - t32:
            p32.dt = t64 - t32
        _, a, u = c32.dev_rk45_step(b32, p32, mat, 2, 1e-5, 0.8, 1e-18, 1e-12)
        if a:
            t32 += u
    m64, m32 = c64.dev_download(b64), c32.dev_download(b32)
    d_avg = np.abs(m32.mean(axis=0) - m64.mean(axis=0)).max()
    d_cell = np.abs(m32 - m64).max()
    bf64 = c64.build_h_eff(m64, L.LLGParams(alpha=0.3, b_ext=bext), mat)
    bf32 = c32.build_h_eff(m64, L.LLGParams(alpha=0.3, b_ext=bext), mat)
    e_b = np.sqrt(((bf32 - bf64) ** 2).sum(1)).max() / np.sqrt((bf64 ** 2).sum(1)).max()
    print(f"C2 after 200 steps ({t64 * 1e12:.3f} ps): <m> {d_avg:.2e}, max cell {d_cell:.2e}, field {e_b:.2e}")
    assert d_avg <= 1e-6 and d_cell <= 1e-4 and e_b <= 1e-5


def _run_vs_steps(L, films, P, es, attempts, split, tail=0, single=False):
    """Films stepped by device-controlled runs (attempt counts `split`) and
    then `tail` single attempts, against the same films stepped `attempts`
    times one attempt per call with the host controller."""
    from conftest import DeviceBuffer
    import copy
    ctx_r = [L.Context(g, precision=P) for g, _, _, _ in films]
    ctx_s = [L.Context(g, precision=P) for g, _, _, _ in films]
    buf_r = [DeviceBuffer(3 * g.n_cells() * es) for g, _, _, _ in films]
    buf_s = [DeviceBuffer(3 * g.n_cells() * es) for g, _, _, _ in films]
    for (g, _, _, m0), cr, cs, br, bs in zip(films, ctx_r, ctx_s, buf_r, buf_s):
        cr.dev_upload(m0, br)
        cs.dev_upload(m0, bs)
    pr = [copy.deepcopy(f[2]) for f in films]
    ps = [copy.deepcopy(f[2]) for f in films]
    pr[1].dt = ps[1].dt = 5e-11          # one film starts with a step that is rejected
    mats = [f[1] for f in films]
    out_r = []
    for k in split:
        if single:   # llgmi_dev_rk45_run film by film (the engine's own kernels)
            per = [c.dev_rk45_run(b, p, m, 2, 1e-5, 0.8, 1e-18, 1e-11, k) for c, b, p, m in zip(ctx_r, buf_r, pr, mats)]
            out_r += [[per[i][a] for i in range(len(films))] for a in range(k)]
        else:
            out_r += L.dev_rk45_run_batch(ctx_r, buf_r, pr, mats, k, 2, 1e-5, 0.8, 1e-18, 1e-11)
    for _ in range(tail):
        out_r.append([c.dev_rk45_step(b, p, m, 2, 1e-5, 0.8, 1e-18, 1e-11)
                      for c, b, p, m in zip(ctx_r, buf_r, pr, mats)])
    out_s = [[c.dev_rk45_step(b, p, m, 2, 1e-5, 0.8, 1e-18, 1e-11) for c, b, p, m in zip(ctx_s, buf_s, ps, mats)]
             for _ in range(attempts)]
    states = [(cr.dev_download(br), cs.dev_download(bs)) for cr, cs, br, bs in zip(ctx_r, ctx_s, buf_r, buf_s)]
    return out_r, out_s, pr, ps, states


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rk45_device_controlled_run_matches_attempts_one_at_a_time(gpu_lib, prec, single):
    """llgmi_dev_rk45_run_batch (controller on the device, 8 attempts in one
    call) against llgmi_dev_rk45_step per film and attempt: the same accept
    decisions; step sizes, eps and states close.  The device's pow rounds
    differently from the host libm's in the last ulp, and a step size an ulp
    apart moves the next eps by up to ~1e-10 relative (eps is dt times a sum
    that cancels to ~1e-4 of the slopes), so step sizes are compared at 1e-9
    relative and eps at 1e-6 (FP64; FP32 stage coefficients round to FP32, so
    1e-3 and 1e-2); the first attempt of one film is rejected."""
    L = gpu_lib
    P = L.F64 if prec == "f64" else L.F32
    es = 8 if prec == "f64" else 4
    print(f"device-controlled RK45 run vs single attempts {prec} single={single}", flush=True)
    out_r, out_s, pr, ps, states = _run_vs_steps(L, _films(3, 6), P, es, 8, [8], single=single)
    tol_dt, tol_eps, tol_m = (1e-9, 1e-6, 1e-12) if prec == "f64" else (1e-3, 1e-2, 1e-5)
    assert not out_r[0][1][1] and not out_s[0][1][1]     # the large first step is rejected
    for a, (rr, rs) in enumerate(zip(out_r, out_s)):
        for i, ((er, ar, ur), (es_, as_, us)) in enumerate(zip(rr, rs)):
            assert ar == as_, (a, i)
            assert abs(ur - us) <= tol_dt * us, (a, i, ur, us)
            assert abs(er - es_) <= tol_eps * abs(es_) + 1e-30, (a, i, er, es_)
    for i, (a, b) in enumerate(zip(pr, ps)):
        assert abs(a.dt - b.dt) <= tol_dt * b.dt, i
    for i, (mr, ms) in enumerate(states):
        assert np.abs(mr - ms).max() <= tol_m, i


@pytest.mark.parametrize("single", [False, True])
def test_rk45_device_controlled_runs_chain_with_single_steps(gpu_lib, single):
    """Runs of 3 and 2 attempts, then single attempts on the same contexts:
    the FSAL state and k1/k7 the runs leave behind carry over (FP64, against
    8 single attempts)."""
    L = gpu_lib
    print(f"device-controlled RK45 runs chained with single attempts single={single}", flush=True)
    out_r, out_s, pr, ps, states = _run_vs_steps(L, _films(7, 4), L.F64, 8, 8, [3, 2], tail=3, single=single)
    assert len(out_r) == 8
    for a, (rr, rs) in enumerate(zip(out_r, out_s)):
        for i, ((er, ar, ur), (es_, as_, us)) in enumerate(zip(rr, rs)):
            assert ar == as_ and abs(ur - us) <= 1e-9 * us, (a, i)
            assert abs(er - es_) <= 1e-6 * abs(es_) + 1e-30, (a, i)
    for i, (mr, ms) in enumerate(states):
        assert np.abs(mr - ms).max() <= 1e-12, i


def test_rk45_device_controlled_run_needs_the_fft_demag(gpu_lib):
    from conftest import DeviceBuffer
    L = gpu_lib
    g = L.Grid2D(16, 16, 3e-9, 3e-9, 3e-9)
    c = L.Context(g, precision=L.F64)
    b = DeviceBuffer(3 * g.n_cells() * 8)
    c.dev_upload(np.tile([0.0, 0.0, 1.0], (g.n_cells(), 1)), b)
    p = L.LLGParams(alpha=0.1, dt=1e-13)
    mat = L.Material(ms=8e5, a_ex=1.3e-11, demag=True, demag_method=L.DemagMethod.PoissonMG)
    with pytest.raises(L.LlgmiError) as e:
        c.dev_rk45_run(b, p, mat, 2, 1e-5, 0.8, 1e-18, 1e-11, 4)
    assert e.value.code == L.ERR_UNSUPPORTED
    assert p.dt == 1e-13
    with pytest.raises(L.LlgmiError) as e:
        c.dev_rk45_run(b, p, L.Material(ms=8e5, a_ex=1.3e-11, demag=True), 2, 1e-5, 0.8, 1e-18, 1e-11, 257)
    assert e.value.code == L.ERR_INVALID
