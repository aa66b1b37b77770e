"""Benchmark: field evaluations/s of the uniform-grid LLG hot path on one MI355X.

Workload (1024 x 1024 x 1 film, FP32): dx = dy = 2 nm, dz = 1 nm, Ms = 5.8e5 A/m,
A = 1.5e-11 J/m, Ku = 8e5 J/m^3 along z, interfacial D = 3e-3 J/m^2,
alpha = 0.3, B_ext = (-24.6, 4.3, 0) mT, FFT demag zero-padded to 2048 x 2048,
m0 = seeded +-z Voronoi domains (2048 points) with 10 % angular noise.

One step = one full field evaluation (build_h_eff_masked, FieldMask::Full:
Zeeman + exchange + anisotropy + DMI + FFT demag) with m already resident in
HBM, called through the C ABI (llgmi_dev_build_h_eff).

The timed region holds K evaluations and nothing else.  The JSON line also
carries:
  per_eval      median and p10/p90 of per-evaluation durations (a second pass
                of K evaluations with one HIP event after each);
  roofline      the dominant kernel (column FFT * tensor * inverse column FFT):
                its average duration from launches back to back between two
                HIP events on the library stream, algorithmic bytes per launch
                as below;
  cpu_baseline  the CPU oracle (C FP64 restatement of the reference path,
                OpenMP over the host cores, and one thread) on the same 1024^2
                workload, 5 evaluations each after an untimed warm-up that
                builds its kernel;
  fp64_gpu      the same evaluation in an FP64 context;
  rel_err_vs_fp64  FP32 against the FP64 GPU path on the same m, per term;
  kernels_us_event_timed  per-kernel durations from a separate pass with events
                around every launch (includes dispatch latency);
  rk45          accepted RK45 steps/s (FSAL, 6 field evaluations per step);
  rk23_relax    accepted relax-mode RK23 steps/s (FSAL, 3 field evaluations per
                step) and energy breakdowns/s;
  host_aos      the same evaluation through the host entry point (AoS f64 in and
                out, PCIe transfers included);
  batched_films 64 independent films of mixed sizes (the batched workload),
                one context and stream per film, evaluated side by side, and
                per-film adaptive RK45 through the batched stepper;
  mg            the multigrid Poisson demag on a 512 x 512 film with a square
                lattice of holes (3 nm cells, 7-point operator, red-black SOR,
                injection, rediscretised coarse levels, treecode boundary,
                tol_rel 1e-8): cold and warm-started solves on the GPU and the
                CPU oracle's time for the same solves.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "llg-sim_amd"))

METRIC = "field evals/s at 1024² FP32 vs FP64 plain (rel err ≤1e-5); % HBM peak"
HBM_PEAK_GBS = 8000.0

NX = NY = 1024
CELL = (2e-9, 2e-9, 1e-9)
MAT = dict(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, easy_axis=(0.0, 0.0, 1.0), dmi=3e-3, demag=True)
BEXT = (-24.6e-3, 4.3e-3, 0.0)
ALPHA = 0.3


# HBM bytes per launch of the column pass at this workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE counters (profiles/field_eval_1024_fp32_pmc_traffic.md),
# FETCH_SIZE scaled by 2.0 (a plain 256 MiB copy reads as 128 MiB).
COL_TRAFFIC_BYTES = 67_780_000


def col_pass_bytes(nx, ny, elem):
    """Algorithmic bytes of one column-pass launch: read + write the half
    spectrum (3 components x px/2 complex columns x ny rows: the real DC and
    Nyquist columns share one complex column) and read the real
    quadrant-symmetric kernel ((px/2+1) x (py/2+1) x 4 reals)."""
    px, py = 2 * nx, 2 * ny
    hx, hy = px // 2 + 1, py // 2 + 1
    return 2 * 3 * (px // 2) * ny * 2 * elem + hx * hy * 4 * elem


def field_eval_bytes(nx, ny, elem):
    """Algorithmic bytes of the whole pipeline per field evaluation."""
    n = nx * ny
    spec = 3 * nx * ny * 2 * elem
    row_fwd = 3 * n * elem + spec
    row_inv = spec + 3 * n * elem
    local = 3 * 3 * n * elem
    return row_fwd + col_pass_bytes(nx, ny, elem) + row_inv + local


def rel_err(a, b):
    import numpy as np
    a = np.asarray(a).reshape(-1, 3)
    b = np.asarray(b).reshape(-1, 3)
    den = np.sqrt((b * b).sum(1)).max()
    return float(np.sqrt(((a - b) ** 2).sum(1)).max() / den) if den > 0 else 0.0


def accuracy_report(L, m_host):
    """FP32 vs FP64 GPU path on the same m, total and per field term."""
    grid = L.Grid2D(NX, NY, *CELL)
    c32 = L.Context(grid, precision=L.F32)
    c64 = L.Context(grid, precision=L.F64)
    p = L.LLGParams(alpha=ALPHA, b_ext=BEXT)
    out = {}
    full = L.Material(**MAT)
    out["total"] = rel_err(c32.build_h_eff(m_host, p, full), c64.build_h_eff(m_host, p, full))
    p0 = L.LLGParams(alpha=ALPHA)
    ex = L.Material(ms=MAT["ms"], a_ex=MAT["a_ex"])
    out["exchange"] = rel_err(c32.build_h_eff(m_host, p0, ex), c64.build_h_eff(m_host, p0, ex))
    an = L.Material(ms=MAT["ms"], k_u=MAT["k_u"], easy_axis=MAT["easy_axis"])
    out["anisotropy"] = rel_err(c32.build_h_eff(m_host, p0, an), c64.build_h_eff(m_host, p0, an))
    dm = L.Material(ms=MAT["ms"], a_ex=MAT["a_ex"], dmi=MAT["dmi"])
    d32 = c32.build_h_eff_masked(m_host, p0, dm, L.FieldMask.ExchAnisDmi) - c32.build_h_eff_masked(m_host, p0, dm, L.FieldMask.ExchAnis)
    d64 = c64.build_h_eff_masked(m_host, p0, dm, L.FieldMask.ExchAnisDmi) - c64.build_h_eff_masked(m_host, p0, dm, L.FieldMask.ExchAnis)
    out["dmi"] = rel_err(d32, d64)
    out["demag"] = rel_err(c32.compute_demag_field(m_host, full), c64.compute_demag_field(m_host, full))
    bx = np.array(BEXT)
    out["zeeman"] = float(np.linalg.norm(bx.astype(np.float32).astype(np.float64) - bx) / np.linalg.norm(bx))
    c32.close()
    c64.close()
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m_host, n_eval=5):
    """CPU oracle (C FP64 restatement of the reference path: C2C FFTs of the
    2nx x 2ny padded grid, serial stencils, OpenMP where the reference uses
    rayon) on this workload, at the host's OpenMP thread count and at one
    thread; median and p10/p90 of n_eval evaluations after an untimed warm-up
    that builds the kernel."""
    from oracle import oracle as O
    threads = O.get_threads()
    ctx = O.Context(O.Grid(NX, NY, *CELL))
    p = O.Params(alpha=ALPHA, b_ext=BEXT)
    mat = O.Material(**MAT)
    t0 = time.perf_counter()
    ctx.build_h_eff(m_host, p, mat, O.FULL)  # warm-up: builds the kernel, untimed
    t_build = time.perf_counter() - t0

    def timed(n):
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            ctx.build_h_eff(m_host, p, mat, O.FULL)
            ts.append(time.perf_counter() - t0)
        return np.array(ts)

    t_all = timed(n_eval)
    O.set_threads(1)
    try:
        t_one = timed(n_eval)
    finally:
        O.set_threads(threads)
    med = float(np.median(t_all))
    return {"value": round(1.0 / med, 3), "unit": "field evals/s", "cores": threads, "kind": "port",
            "sample": f"{n_eval} full field evaluations (1024x1024 FP64, C2C FFT demag 2048^2) at {threads} "
                      f"threads and {n_eval} at 1 thread, after one untimed warm-up that built the kernel "
                      f"({t_build:.1f} s); value = 1 / median",
            "cpu_model": _cpu_model(),
            "ms_median": round(1e3 * med, 1), "ms_p10": round(1e3 * float(np.percentile(t_all, 10)), 1),
            "ms_p90": round(1e3 * float(np.percentile(t_all, 90)), 1),
            "one_thread": {"value": round(1.0 / float(np.median(t_one)), 4),
                           "ms_median": round(1e3 * float(np.median(t_one)), 1),
                           "ms_p10": round(1e3 * float(np.percentile(t_one, 10)), 1),
                           "ms_p90": round(1e3 * float(np.percentile(t_one, 90)), 1)}}


def fp64_rate(L, m_host, n_eval):
    """The same field evaluation in an FP64 context (device buffers, state
    resident): evaluations/s, the like-for-like figure for the FP64 CPU path."""
    import torch
    ctx = L.Context(L.Grid2D(NX, NY, *CELL), precision=L.F64)
    m = torch.empty((3, NY, NX), dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(m)
    ctx.dev_upload(m_host, m)
    params = L.LLGParams(alpha=ALPHA, b_ext=BEXT)
    mat = L.Material(**MAT)
    for _ in range(5):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_eval):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    passes = ctx.profile_fft_passes(m, MAT["ms"], 50)
    ctx.close()
    return {"value": round(n_eval / dt, 1), "unit": "field evals/s", "evals": n_eval,
            "fft_passes_back_to_back_us": {k: round(v, 2) for k, v in passes.items()},
            "col_pass_frac_hbm": round(col_pass_bytes(NX, NY, 8) / (passes["fft_col_mul"] * 1e-6) / 1e9
                                       / HBM_PEAK_GBS, 4)}


MG_N = 512
MG_CELL = (3e-9, 3e-9, 3e-9)
MG_MS = 8.6e5


def mg_workload():
    """512 x 512 film, 3 nm cells, holes of radius 24 nm on a 96 nm square
    lattice (m = 0 there), seeded smoothed-random m elsewhere."""
    from initial_states import hole_lattice_mask, smoothed_random
    mask = hole_lattice_mask(MG_N, MG_N, MG_CELL[0], MG_CELL[1], pitch=96e-9, radius=24e-9)
    m = smoothed_random(MG_N, MG_N, seed=11).reshape(-1, 3)
    m[~mask.reshape(-1)] = 0.0
    return m


MG_CFG = dict(stencil=0, prolong=0, coarse_op=0, smoother=1, bc=2, tol_rel=1e-8, v_cycles_max=300)


def mg_bench(L, cpu: bool, reps: int = 5):
    """Cold (phi = 0) and warm-started solves of the multigrid demag."""
    m = mg_workload()
    ctx = L.Context(L.Grid2D(MG_N, MG_N, *MG_CELL), precision=L.F64)
    mat = L.Material(ms=MG_MS, demag=True)
    out = {"workload": "512x512 film, 3 nm cells, hole lattice (pitch 96 nm, r 24 nm), box 524x524x36, "
                       "7-point / red-black SOR / injection / rediscretised / treecode, tol_rel 1e-8"}
    for warm in (0, 1):
        ctx.set_mg_config(L.mg_config(warm_start=warm, **MG_CFG))
        ctx.compute_demag_field_poisson_mg(m, mat)          # untimed: hierarchy, tree, first solve
        ctx.profile_reset()
        ctx.profile_enable([L.K_MG, L.K_MG_BC])
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.compute_demag_field_poisson_mg(m, mat)
        dt = (time.perf_counter() - t0) / reps
        ctx.profile_enable([])
        st = ctx.mg_last_stats()
        bc_ms, bc_n = ctx.profile_read(L.K_MG_BC)
        key = "warm" if warm else "cold"
        out[key] = {"solve_ms": round(1e3 * dt, 3), "v_cycles": st["cycles"],
                    "boundary_ms": round(bc_ms / max(bc_n, 1), 3),
                    "ms_per_v_cycle": round((1e3 * dt - bc_ms / max(bc_n, 1)) / max(st["cycles"], 1), 3),
                    "final_residual_over_tol": round(st["last_residual"] / st["tol_target"], 4)}
    ctx.close()
    if cpu:
        from oracle import oracle as O
        g = O.Grid(MG_N, MG_N, *MG_CELL)
        mg = O.MG(g, O.mg_config(warm_start=1, **MG_CFG))
        b = np.zeros_like(m)
        t0 = time.perf_counter()
        mg.add_field(m, b, MG_MS)
        t_cold = time.perf_counter() - t0
        t0 = time.perf_counter()
        mg.add_field(m, b, MG_MS)
        t_warm = time.perf_counter() - t0
        out["cpu_oracle"] = {"cold_solve_s": round(t_cold, 3), "warm_solve_s": round(t_warm, 3),
                             "cores": O.get_threads(), "kind": "port",
                             "sample": "one cold and one warm-started solve of the same state"}
        out["speedup_vs_cpu"] = {"cold": round(t_cold / (out["cold"]["solve_ms"] * 1e-3), 1),
                                 "warm": round(t_warm / (out["warm"]["solve_ms"] * 1e-3), 1)}
    return out


def films_bench(L, n_films: int = 64, rounds: int = 20, rk_rounds: int = 10):
    """Independent films: one context per film, sizes from
    {64,128,256,512} x {64,128,256}, dx in [2, 5] nm, dz = 3 nm, random
    materials, smoothed random m.  Field evaluations: a round queues one full
    field evaluation (Full mask, FFT demag) on every film's own stream before
    one synchronisation.  RK45: one llgmi_dev_rk45_step_batch attempt of all
    films per round (one launch per kernel class over all films, replayed from
    one HIP graph), against the same films stepped one after another."""
    import torch
    from initial_states import batched_films
    specs = batched_films(64, n_films)
    films = []
    for f in specs:
        nx, ny = f["nx"], f["ny"]
        ctx = L.Context(L.Grid2D(nx, ny, *f["cell"]), precision=L.F32)
        m = torch.empty((3, ny, nx), dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(m)
        ctx.dev_upload(f["m0"], m)
        films.append((ctx, m, b, L.Material(**f["material"]), nx * ny))
    params = L.LLGParams(alpha=0.02, b_ext=(0.01, 0.0, 0.0))
    for ctx, m, b, mat, _ in films:          # untimed: kernels, autotune
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    for ctx, *_ in films:
        ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(rounds):
        for ctx, m, b, mat, _ in films:
            ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    for ctx, *_ in films:
        ctx.synchronize()
    dt = time.perf_counter() - t0
    cells = sum(f[4] for f in films)
    out = {"workload": f"{n_films} films, sizes {{64,128,256,512}}x{{64,128,256}}, FP32, field evaluations "
                       "queued film by film on one stream per film",
           "films": n_films, "total_cells": cells, "rounds": rounds,
           "film_evals_per_s": round(n_films * rounds / dt, 1),
           "cell_updates_per_s": round(cells * rounds / dt, 1),
           "ms_per_round": round(1e3 * dt / rounds, 3)}
    # per-film adaptive RK45 (own alpha and dt per film), one batched attempt
    # per round (llgmi_dev_rk45_step_batch) against the same films stepped one
    # after another (llgmi_dev_rk45_step)
    ctxs = [f[0] for f in films]
    ms = [f[1] for f in films]
    mats = [f[3] for f in films]
    pars = [L.LLGParams(alpha=f["alpha"], dt=1e-13, b_ext=(0.01, 0.0, 0.0)) for f in specs]
    rk = dict(mask=L.FieldMask.Full, max_err=1e-5, headroom=0.8, dt_min=1e-18, dt_max=1e-11)
    for _ in range(6):       # untimed: the eager first attempts and the captures of both FSAL parities
        L.dev_rk45_step_batch(ctxs, ms, pars, mats, **rk)
    acc = att = 0
    t0 = time.perf_counter()
    for _ in range(rk_rounds):
        for e, a, u in L.dev_rk45_step_batch(ctxs, ms, pars, mats, **rk):
            acc += a
            att += 1
    dtb = time.perf_counter() - t0
    acc_s = 0
    t0 = time.perf_counter()
    for _ in range(rk_rounds):
        for c, m, p, mat in zip(ctxs, ms, pars, mats):
            acc_s += c.dev_rk45_step(m, p, mat, **rk)[1]
    dts = time.perf_counter() - t0
    # the controller on the device: rk_rounds attempts of every film per call
    L.dev_rk45_run_batch(ctxs, ms, pars, mats, rk_rounds, **rk)          # untimed: graph capture
    t0 = time.perf_counter()
    res_d = L.dev_rk45_run_batch(ctxs, ms, pars, mats, rk_rounds, **rk)
    dtd = time.perf_counter() - t0
    acc_d = sum(int(a) for row in res_d for _, a, _ in row)
    out["rk45"] = {"workload": "per-film adaptive Dormand-Prince RK45, max_err 1e-5, alpha in U[0.01, 0.5], FP32 "
                               "state and fields, FP64 error norm and controller; batched kernels over all "
                               "films, one graph replay and one host synchronisation per round",
                   "rounds": rk_rounds, "attempts": att, "accepted": acc,
                   "accepted_steps_per_s": round(acc / dtb, 1),
                   "cell_steps_per_s": round(acc / att * cells * rk_rounds / dtb, 1),
                   "ms_per_round": round(1e3 * dtb / rk_rounds, 3),
                   "one_film_at_a_time_steps_per_s": round(acc_s / dts, 1),
                   "device_controller": {"accepted_steps_per_s": round(acc_d / dtd, 1),
                                         "attempts": len(films) * rk_rounds, "accepted": acc_d,
                                         "ms_per_round": round(1e3 * dtd / rk_rounds, 3),
                                         "attempts_per_call": rk_rounds}}
    # per-film accuracy: the FP32 field of each film's current state against an
    # FP64 context on the same state (max |dB| / max |B| over the film's cells)
    errs = []
    for (ctx, m, b, mat, _), f in zip(films, specs):
        mh = ctx.dev_download(m)
        p = L.LLGParams(alpha=f["alpha"], b_ext=(0.01, 0.0, 0.0))
        c64 = L.Context(L.Grid2D(f["nx"], f["ny"], *f["cell"]), precision=L.F64)
        errs.append(rel_err(ctx.build_h_eff(mh, p, mat), c64.build_h_eff(mh, p, mat)))
        c64.close()
    out["field_rel_err_vs_fp64"] = {"max": float(max(errs)), "median": float(np.median(errs)),
                                    "films_over_1e-5": int(sum(e > 1e-5 for e in errs))}
    for ctx, *_ in films:
        ctx.close()
    return out


def warm_start_bench():
    """The learned multigrid initial guess on the 512 x 512 film with holes:
    V-cycles to tol_rel 1e-6 from the zero guess, the previous state's
    potential, the network's potential, the previous potential plus the
    network's change, and the previous potential plus the difference
    network's potential of m - m_prev, along a relaxation the networks were
    not trained on
    (states 1 and 20 relaxation steps apart), with the network's MFMA rates.
    Uses the committed weights (llg-sim_amd/weights); they are trained by
    llg-sim_amd/warmstart.py."""
    import warmstart as WS
    if not os.path.exists(WS.DEFAULT_WEIGHTS):
        return {"skipped": "no trained weights"}
    params = np.load(WS.DEFAULT_WEIGHTS)
    info = os.path.splitext(WS.DEFAULT_WEIGHTS)[0] + ".json"
    prob = WS.Problem()
    out = {"weights": os.path.relpath(WS.DEFAULT_WEIGHTS, REPO)}
    if os.path.exists(info):
        with open(info) as f:
            d = json.load(f)
        out["training"] = {k: d[k] for k in ("pairs", "train_steps", "batch", "record_s", "train_s",
                                             "loss_first_100", "loss_last_100") if k in d}
    p_delta = np.load(WS.DELTA_WEIGHTS) if os.path.exists(WS.DELTA_WEIGHTS) else None
    names = ("zero", "previous", "learned", "previous_plus_learned_change") + \
        (("previous_plus_learned_delta",) if p_delta is not None else ())
    for stride, states in ((1, 20), (20, 10)):
        c = WS.compare(prob, params, states=states, stride=stride, params_delta=p_delta)
        out[f"stride_{stride}"] = {"workload": c["workload"], "inference_ms": c["inference_ms"],
                                   "max_dm_between_states": c["max_dm_between_states_mean"],
                                   **{f"{n}_v_cycles": c[n]["v_cycles_mean"] for n in names},
                                   **{f"{n}_ms": c[n]["ms_mean_incl_inference"] for n in names}}
    if p_delta is not None:
        out["guess_error_study"] = {f"stride_{k}": WS.guess_error_study(prob, p_delta, stride=k, states=n)
                                    for k, n in ((1, 10), (20, 5))}
    out["mfma"] = WS.measure_throughput(prob, params)
    return out


def rk45_rate(L, ctx, m, mat, steps):
    """Accepted RK45 steps/s at the headline workload (FSAL, state resident,
    max_err 1e-5), after six untimed attempts (eager first attempts, then the
    graph captures of both FSAL parities)."""
    mr = m.clone()
    rp = L.LLGParams(alpha=ALPHA, dt=1e-14, b_ext=BEXT)
    for _ in range(6):
        ctx.dev_rk45_step(mr, rp, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12)
    ctx.synchronize()
    t0 = time.perf_counter()
    acc = att = 0
    while acc < steps and att < 4 * steps:
        _, a, _ = ctx.dev_rk45_step(mr, rp, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12)
        acc += int(a)
        att += 1
    ctx.synchronize()
    dtr = time.perf_counter() - t0
    out = {"accepted_steps_per_s": round(acc / dtr, 1), "attempts": att, "accepted": acc,
           "us_per_attempt": round(1e6 * dtr / att, 1), "max_err": 1e-5, "last_dt": rp.dt}
    # the same stepping with the controller on the device: `steps` attempts
    # per llgmi_dev_rk45_run call (one graph replay, one host synchronisation)
    rr = L.LLGParams(alpha=ALPHA, dt=rp.dt, b_ext=BEXT)
    ctx.dev_rk45_run(mr, rr, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12, steps)   # untimed: capture
    ctx.synchronize()
    t0 = time.perf_counter()
    res = []
    for _ in range(3):
        res += ctx.dev_rk45_run(mr, rr, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12, steps)
    dtd = time.perf_counter() - t0
    acc_d = sum(int(a) for _, a, _ in res)
    out["device_controller"] = {"accepted_steps_per_s": round(acc_d / dtd, 1), "attempts": len(res),
                                "accepted": acc_d, "us_per_attempt": round(1e6 * dtd / len(res), 1),
                                "attempts_per_call": steps}
    return out


def host_aos_rate(L, m_host, n=20):
    """Field evaluations/s through the host entry point (AoS f64 in and out)."""
    ctx = L.Context(L.Grid2D(NX, NY, *CELL), precision=L.F32)
    p = L.LLGParams(alpha=ALPHA, b_ext=BEXT)
    mat = L.Material(**MAT)
    ctx.build_h_eff(m_host, p, mat)
    t0 = time.perf_counter()
    for _ in range(n):
        ctx.build_h_eff(m_host, p, mat)
    dt = time.perf_counter() - t0
    ctx.close()
    return round(n / dt, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--rk45-steps", type=int, default=20)
    ap.add_argument("--quick", action="store_true",
                    help="field-evaluation loop only (no accuracy check, CPU leg or RK45); for profilers")
    ap.add_argument("--no-mg", action="store_true", help="skip the multigrid leg")
    ap.add_argument("--mg-only", action="store_true",
                    help="run only the multigrid leg (GPU solves, no CPU oracle); for profilers")
    ap.add_argument("--ws-only", action="store_true",
                    help="run only the learned initial guess's forward / training-step rates; for profilers")
    ap.add_argument("--rk45-only", action="store_true",
                    help="run only the RK45 leg at the headline workload (--rk45-steps steps); for profilers")
    ap.add_argument("--films-only", action="store_true",
                    help="run only the batched-films leg (64 films); for profilers")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of 256 MiB (x10) to scale HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU slab "
              "decomposition is a design, not a build", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4,
                          "launches": 10}), flush=True)
        return
    if args.quick:
        args.no_cpu = True
        args.no_mg = True
        args.rk45_steps = 0

    global np
    import numpy as np
    import torch
    import llgmi as L
    from initial_states import voronoi_domains

    if args.mg_only:
        print(json.dumps({"mg": mg_bench(L, cpu=False)}), flush=True)
        return
    if args.films_only:
        print(json.dumps({"batched_films": films_bench(L)}), flush=True)
        return
    if args.ws_only:
        import warmstart as WS
        prob = WS.Problem()
        print(json.dumps({"mfma": WS.measure_throughput(prob, np.load(WS.DEFAULT_WEIGHTS))}), flush=True)
        return

    assert torch.cuda.is_available(), "no GPU visible to torch"
    m_host = voronoi_domains(NX, NY, seed=2048, n_points=2048, noise=0.1).reshape(-1, 3)
    print(f"[bench] workload 1024x1024 FP32, steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)

    grid = L.Grid2D(NX, NY, *CELL)
    ctx = L.Context(grid, precision=L.F32)
    dev = torch.device("cuda:0")
    m = torch.empty((3, NY, NX), dtype=torch.float32, device=dev)
    b = torch.empty((3, NY, NX), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ctx.dev_upload(m_host, m)
    params = L.LLGParams(alpha=ALPHA, b_ext=BEXT)
    mat = L.Material(**MAT)
    if args.rk45_only:
        print(json.dumps({"rk45": rk45_rate(L, ctx, m, mat, args.rk45_steps)}), flush=True)
        return

    for _ in range(args.warmup):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    ctx.synchronize()

    # timed region: K evaluations, nothing else on the stream
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = args.steps / elapsed

    # per-evaluation durations: one event after each evaluation on the
    # library stream (a second pass of K evaluations)
    lib_stream = torch.cuda.ExternalStream(ctx.stream(), device=dev)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(lib_stream)
    for i in range(args.steps):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
        evs[i + 1].record(lib_stream)
    ctx.synchronize()
    per_eval_us = np.array([1e3 * evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)])
    dist = {"median_us": round(float(np.median(per_eval_us)), 2),
            "p10_us": round(float(np.percentile(per_eval_us, 10)), 2),
            "p90_us": round(float(np.percentile(per_eval_us, 90)), 2), "evals": args.steps}

    # the roofline kernel: each FFT pass launched back to back between two
    # HIP events on the library stream (no event between launches)
    reps = max(100, args.steps)
    passes_us = ctx.profile_fft_passes(m, MAT["ms"], reps)
    col_us = passes_us["fft_col_mul"]
    col_bytes = col_pass_bytes(NX, NY, 4)
    achieved = col_bytes / (col_us * 1e-6) / 1e9

    # per-kernel breakdown inside the pipeline (separate pass, every launch
    # between its own pair of events, which adds the dispatch latency)
    ctx.profile_reset()
    ctx.profile_enable(None)
    for _ in range(50):
        ctx.dev_build_h_eff(m, b, params, mat, L.FieldMask.Full)
    ctx.synchronize()
    ctx.profile_enable([])
    kernels = {}
    for cls, name in L.KERNEL_CLASS_NAMES.items():
        t, c = ctx.profile_read(cls)
        if c:
            kernels[name] = round(1e3 * t / c, 2)

    fp64 = None
    if not args.quick:
        fp64 = fp64_rate(L, m_host, max(20, args.steps // 5))
        print(f"[bench] fp64: {fp64}", file=sys.stderr, flush=True)

    # RK45 accepted steps/s (FSAL; state resident)
    rk = rk45_rate(L, ctx, m, mat, args.rk45_steps) if args.rk45_steps > 0 else {}

    # relax-mode RK23 accepted steps/s (FSAL, 3 field evaluations per step)
    # and energy breakdowns/s (one demag field + one reduction each)
    r23 = {}
    if args.rk45_steps > 0:
        mr = m.clone()
        rp = L.LLGParams(alpha=ALPHA, dt=1e-14, b_ext=BEXT)
        for _ in range(3):
            ctx.dev_rk23_relax_step(mr, rp, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12)
        ctx.synchronize()
        t0 = time.perf_counter()
        acc = att = 0
        while acc < args.rk45_steps and att < 4 * args.rk45_steps:
            _, a, _ = ctx.dev_rk23_relax_step(mr, rp, mat, L.FieldMask.Full, 1e-5, 0.8, 1e-18, 1e-12)
            acc += int(a)
            att += 1
        ctx.synchronize()
        dtr = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(10):
            ctx.dev_compute_energy_geom(mr, mat, BEXT)
        dte = time.perf_counter() - t0
        r23 = {"accepted_steps_per_s": round(acc / dtr, 1), "attempts": att, "accepted": acc,
               "max_err": 1e-5, "last_dt": rp.dt, "energy_evals_per_s": round(10 / dte, 1)}

    acc_rep = None
    host_rate = None
    if not args.quick:
        acc_rep = accuracy_report(L, m_host)
        print(f"[bench] accuracy done: {acc_rep}", file=sys.stderr, flush=True)
        host_rate = host_aos_rate(L, m_host)
    films = None
    if not args.quick:
        films = films_bench(L)
        print(f"[bench] films: {films}", file=sys.stderr, flush=True)
    mg = None
    if not args.no_mg:
        mg = mg_bench(L, cpu=not args.no_cpu)
        print(f"[bench] multigrid: {mg}", file=sys.stderr, flush=True)
    ws = None
    if not args.no_mg and not args.quick:
        ws = warm_start_bench()
        print(f"[bench] mg warm start: {ws}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m_host)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "field evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "uniform film 1024x1024x1, dx=dy=2nm dz=1nm, exchange+anisotropy+"
                               "interfacial DMI+Zeeman+FFT demag (padded 2048^2), Voronoi +-z m0",
                   "nx": NX, "ny": NY, "precision": "fp32", "field_mask": "Full"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": COL_TRAFFIC_BYTES,
                     "kernel": "k_col_demag (column FFT x tensor x inverse column FFT)",
                     "bytes_per_launch": col_bytes, "avg_us": round(col_us, 2),
                     "timing": f"{reps} launches back to back between two HIP events on the library stream"},
        "per_eval": dist,
        "fft_passes_back_to_back_us": {k: round(v, 2) for k, v in passes_us.items()},
        "fp64_gpu": fp64,
        "cpu_baseline": cpu,
        "pipeline_bytes_per_eval": field_eval_bytes(NX, NY, 4),
        "pipeline_frac_hbm": round(field_eval_bytes(NX, NY, 4) * value / 1e9 / HBM_PEAK_GBS, 4),
        "kernels_us_event_timed": kernels,
        "rel_err_vs_fp64": acc_rep,
        "rk45": rk,
        "rk23_relax": r23,
        "host_aos_evals_per_s": host_rate,
        "mg": mg,
        "batched_films": films,
        "mg_warm_start": ws,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
