"""Learned initial guess for the multigrid Poisson demag: recording, training
and the V-cycle comparison (the workflow around llgmi.WarmStartNet).

Workload: a 512 x 512 film of 3 nm cells with a square lattice of circular
holes (m = 0 there), Ms = 8.6e5 A/m, solved by the multigrid demag with the
7-point operator, red-black SOR, injection, rediscretised coarse levels and
the treecode boundary (box 524 x 524 x 36).  The network sees
(m_x, m_y, m_z, mask) on a 528 x 528 plane that holds the box's xy extent and
predicts phi / (Ms dz) on all 36 z planes of the box.

* record(): damping-only RK23 relaxations (the library's relax stepper, FP64,
  FFT demag) of seeded smoothed-random films; after every step the state
  with the holes set to m = 0 (hole pitch drawn from [64, 128] nm, radius
  24 nm) and its multigrid potential, converged to tol_rel 1e-8, are stored
  packed (bf16, in HBM).  The film relaxes without its holes: the stepper's
  stage normalisation turns m = 0 cells into +z (normalize(0) = +z), which
  stalls a relaxation with holes at dt_min; the recorded pairs are exact
  (m, phi(m)) pairs of the holed states all the same.
* train(): Adam on random mini-batches, on the device.
* compare(): along a relaxation the network has not seen (pitch 96 nm),
  V-cycles and wall time for the solve to tol_rel 1e-6 (at least one cycle)
  from the zero guess, from the previous step's potential, from the
  network's potential, and from the previous potential plus the network's
  change between the two states.

The reference has no learned component; its solve starts from zero or from
the previous solve (demag_poisson_mg.rs:1947-2057).  The multigrid solver
still decides convergence: every guess is iterated to the same residual.

Run as a script to train and save weights:
  python llg-sim_amd/warmstart.py --runs 20 --steps 100 --out <file.npy>
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import llgmi as L  # noqa: E402
from initial_states import hole_lattice_mask, smoothed_random  # noqa: E402

NX = NY = 512
CELL = (3e-9, 3e-9, 3e-9)
MS = 8.6e5
A_EX = 1.3e-11
BASE_CHANNELS = 32
SOLVER = dict(stencil=0, prolong=0, coarse_op=0, smoother=1, bc=2, v_cycles_max=300)
TOL_RECORD = 1e-8
TOL_SOLVE = 1e-6
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "mg_warmstart_512_holes.npy")
DELTA_WEIGHTS = os.path.join(HERE, "weights", "mg_warmstart_512_holes_delta.npy")


class Problem:
    """One FP64 context with the multigrid demag and device buffers for a
    state, its hole mask and its potential."""

    def __init__(self, nx: int = NX, ny: int = NY):
        self.nx, self.ny = nx, ny
        self.ctx = L.Context(L.Grid2D(nx, ny, *CELL), precision=L.F64)
        self.mat = L.Material(ms=MS, a_ex=A_EX, demag=True, demag_method=L.DemagMethod.PoissonMG)
        self.mat_dyn = L.Material(ms=MS, a_ex=A_EX, demag=True, demag_method=L.DemagMethod.FftUniform)
        self.m_dyn = L.DeviceArray(3 * nx * ny * 8)
        self.m = L.DeviceArray(3 * nx * ny * 8)
        self.dm = L.DeviceArray(3 * nx * ny * 8)       # a normalised state difference
        self.b = L.DeviceArray(3 * nx * ny * 8)
        self.keep = L.DeviceArray(nx * ny)
        self.film(0, 96e-9)
        self.set_tol(TOL_RECORD, warm=False)
        self.solve()                                   # builds the hierarchy
        self.px, self.py, self.pz, self.offx, self.offy, self.offz = self.ctx.mg_box()
        self.H = self.W = (max(self.px, self.py) + 7) // 8 * 8
        self.phi_bytes = self.px * self.py * self.pz * 8
        self.phi = L.DeviceArray(self.phi_bytes)
        self.scale = MS * CELL[2]

    def set_tol(self, tol: float, warm: bool) -> None:
        self.ctx.set_mg_config(L.mg_config(warm_start=int(warm), tol_rel=tol, v_cycles=1, **SOLVER))

    def solve(self, m=None) -> int:
        """MG demag of the current state (or of m; the solver's phi is the
        initial guess when warm starts are on); returns the V-cycles taken."""
        self.b.zero()
        self.ctx.dev_add_demag_field_pbc(self.m if m is None else m, self.b, self.mat)
        self.ctx.synchronize()
        return int(self.ctx.mg_last_stats()["cycles"])

    def film(self, seed: int, pitch: float, radius: float = 24e-9) -> None:
        mask = hole_lattice_mask(self.nx, self.ny, CELL[0], CELL[1], pitch=pitch, radius=radius)
        m = smoothed_random(self.nx, self.ny, seed=seed).reshape(-1, 3)
        self.ctx.dev_upload(m, self.m_dyn)
        self.keep.upload(mask.reshape(-1).astype(np.uint8))
        self.ctx.rk23_invalidate_last_b_eff()
        self._masked_state()

    def _masked_state(self) -> None:
        self.m.copy_from(self.m_dyn, self.m.nbytes)
        self.ctx.dev_mask_cells(self.m, self.keep)

    def relax_step(self, params: L.LLGParams) -> bool:
        """One damping-only RK23 attempt of the film (FFT demag; the stepper
        renormalises m = 0 cells, so the film relaxes without its holes) and
        the masked state for the multigrid; returns whether it was accepted."""
        _, acc, _ = self.ctx.dev_rk23_relax_step(self.m_dyn, params, self.mat_dyn, L.FieldMask.Full, 1e-5, 0.8,
                                                 1e-18, 1e-11)
        self._masked_state()
        return acc

    def state(self) -> np.ndarray:
        return self.ctx.dev_download(self.m)


def forward_flops(H: int, W: int, base: int, c_pad: int, batch: int = 1) -> float:
    """Multiply-add FLOPs (2 per MAC) of one forward pass: 3x3 convolutions of
    the architecture in csrc/unet.hip (input padded to 16 channels)."""
    c = [base << l for l in range(4)]
    convs = []
    for l in range(4):
        convs += [(l, 16 if l == 0 else c[l - 1], c[l]), (l, c[l], c[l])]
    for l in (2, 1, 0):
        convs += [(l, c[l + 1] + c[l], c[l]), (l, c[l], c[l])]
    convs.append((0, c[0], c_pad))
    return float(sum(2 * 9 * ci * co * (H >> l) * (W >> l) for l, ci, co in convs)) * batch


def measure_throughput(prob: "Problem", params: np.ndarray, batch: int = 8, steps: int = 10) -> dict:
    """Forward (batch 1) and training-step (batch `batch`) rates of the
    network at the problem's plane, as MFMA throughput: a training step
    counts 3x the forward FLOPs (forward, data and weight gradients)."""
    pred = Predictor(prob, params)
    pred.run()
    t0 = time.perf_counter()
    for _ in range(steps):
        pred.net.forward(pred.x, pred.y)
    t_fwd = (time.perf_counter() - t0) / steps
    pred.net.close()
    net = L.WarmStartNet(batch, prob.H, prob.W, prob.pz, BASE_CHANNELS, valid_width=prob.px,
                         valid_height=prob.py)
    net.set_params(params)
    c_pad = net.c_pad
    xb = L.DeviceArray(batch * prob.H * prob.W * 16 * 2)
    tb = L.DeviceArray(batch * prob.H * prob.W * c_pad * 2)
    xb.zero()
    tb.zero()
    net.loss_grad(xb, tb)
    t0 = time.perf_counter()
    for _ in range(steps):
        net.loss_grad(xb, tb)
    t_step = (time.perf_counter() - t0) / steps
    net.close()
    f = forward_flops(prob.H, prob.W, BASE_CHANNELS, c_pad)
    peak = 2.5e15   # MI355X dense bf16 MFMA
    return {"forward_gflop": round(f / 1e9, 2), "forward_ms": round(1e3 * t_fwd, 3),
            "forward_tflops": round(f / t_fwd / 1e12, 1), "forward_frac_bf16_peak": round(f / t_fwd / peak, 4),
            "train_batch": batch, "train_step_ms": round(1e3 * t_step, 3),
            "train_tflops": round(3 * f * batch / t_step / 1e12, 1),
            "train_frac_bf16_peak": round(3 * f * batch / t_step / peak, 4)}


class Dataset:
    """Packed inputs and targets in HBM (bf16), one slot per pair."""

    def __init__(self, prob: Problem, n: int):
        self.n = n
        self.c_pad = (prob.pz + 15) // 16 * 16
        self.sx = prob.H * prob.W * 16 * 2
        self.st = prob.H * prob.W * self.c_pad * 2
        self.x = L.DeviceArray(n * self.sx)
        self.t = L.DeviceArray(n * self.st)


def set_delta(prob: Problem, m_now: np.ndarray, m_prev: np.ndarray) -> float:
    """prob.dm = (m_now - m_prev) / s with s the RMS of m_now - m_prev over
    the components (changes concentrate in walls: the RMS, not the maximum,
    keeps typical inputs of order one); returns s."""
    d = m_now - m_prev
    s = float(np.sqrt(np.mean(d * d)))
    prob.ctx.dev_upload(d / s if s > 0 else d, prob.dm)
    return s


def record(prob: Problem, runs: int, steps: int, seed0: int = 0, pitch_range=(64e-9, 128e-9), log=print,
           delta: bool = False):
    """Training pairs along `runs` relaxations of `steps` steps each: the
    state and its potential, or with delta the normalised change between
    consecutive states and the potential of that change (phi is linear in m
    for a fixed geometry, so phi(m) - phi(m_prev) = phi(m - m_prev))."""
    data = Dataset(prob, runs * steps)
    packer = L.WarmStartNet(1, prob.H, prob.W, prob.pz, 16, valid_width=prob.px, valid_height=prob.py)
    rng = np.random.default_rng(seed0)
    k = 0
    t0 = time.perf_counter()
    for r in range(runs):
        pitch = float(rng.uniform(*pitch_range))
        prob.film(seed0 * 1000 + r + 1, pitch)
        params = L.LLGParams(alpha=0.5, dt=1e-13)
        m_first = prob.state()
        n_acc = 0
        m_prev = m_first
        for _ in range(steps):
            n_acc += prob.relax_step(params)
            src = prob.m
            if delta:
                m_now = prob.state()
                set_delta(prob, m_now, m_prev)
                m_prev = m_now
                src = prob.dm
            prob.set_tol(TOL_RECORD, warm=not delta)
            prob.solve(src)
            prob.ctx.mg_phi_read(prob.phi)
            packer.pack_input(0, src, False, prob.nx, prob.ny, prob.offx, prob.offy, data.x.at(k * data.sx))
            packer.pack_phi(0, prob.phi, prob.px, prob.py, prob.pz, prob.scale, data.t.at(k * data.st))
            k += 1
        dm = float(np.abs(prob.state() - m_first).max())
        log(f"[warmstart] recorded run {r + 1}/{runs} (pitch {pitch * 1e9:.1f} nm): {k} pairs, "
            f"{n_acc}/{steps} accepted, dt {params.dt:.2e} s, max|m_end - m_start| {dm:.3f}, "
            f"{time.perf_counter() - t0:.1f} s")
    packer.close()
    return data


def train(prob: Problem, data: Dataset, steps: int, batch: int = 8, lr: float = 5e-4, seed: int = 1,
          log=print, log_every: int = 100):
    """Adam with a cosine learning-rate decay on random mini-batches;
    returns (net, losses)."""
    net = L.WarmStartNet(batch, prob.H, prob.W, prob.pz, BASE_CHANNELS, valid_width=prob.px,
                         valid_height=prob.py, seed=seed)
    xb, tb = L.DeviceArray(batch * data.sx), L.DeviceArray(batch * data.st)
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    for it in range(steps):
        for j, k in enumerate(rng.integers(0, data.n, batch)):
            xb.copy_from(data.x, data.sx, int(k) * data.sx, j * data.sx)
            tb.copy_from(data.t, data.st, int(k) * data.st, j * data.st)
        a = lr * 0.5 * (1.0 + np.cos(np.pi * it / steps)) + 1e-5
        losses.append(net.train_step(xb, tb, a))
        if (it + 1) % log_every == 0:
            log(f"[warmstart] step {it + 1}/{steps}: loss {np.mean(losses[-log_every:]):.4e}, "
                f"{(time.perf_counter() - t0) / (it + 1) * 1e3:.1f} ms/step")
    return net, losses


class Predictor:
    """Batch-1 inference copy of a trained network."""

    def __init__(self, prob: Problem, params: np.ndarray):
        self.prob = prob
        self.net = L.WarmStartNet(1, prob.H, prob.W, prob.pz, BASE_CHANNELS, valid_width=prob.px,
                                  valid_height=prob.py)
        self.net.set_params(params)
        self.x = L.DeviceArray(prob.H * prob.W * 16 * 2)
        self.y_bytes = prob.H * prob.W * self.net.c_pad * 4
        self.y = L.DeviceArray(self.y_bytes)

    def run(self, m=None) -> None:
        """Network output for the problem's current state (or m) into self.y."""
        p = self.prob
        self.net.pack_input(0, p.m if m is None else m, False, p.nx, p.ny, p.offx, p.offy, self.x)
        self.net.forward(self.x, self.y)

    def potential(self, y, phi, sign: float = 1.0, accumulate: bool = False) -> None:
        p = self.prob
        self.net.unpack_phi(0, y, p.px, p.py, p.pz, sign * p.scale, phi, accumulate)


def compare(prob: Problem, params: np.ndarray, states: int = 30, seed: int = 424242, pitch: float = 96e-9,
            skip: int = 5, stride: int = 1, params_delta=None):
    """V-cycles and wall time to tol_rel 1e-6 from the initial guesses below,
    along a relaxation the networks were not trained on; consecutive compared
    states are `stride` relaxation steps apart (the previous-step guess is
    the potential of the state `stride` steps earlier).  With params_delta,
    also phi_prev + s N_delta((m - m_prev) / s).  The difference m - m_prev is
    formed on the host in this harness and is not in the timed figures."""
    pred = Predictor(prob, params)
    pdel = Predictor(prob, params_delta) if params_delta is not None else None
    prob.film(seed, pitch)
    p_llg = L.LLGParams(alpha=0.5, dt=1e-13)
    for _ in range(skip):
        prob.relax_step(p_llg)
    phi_prev, guess = L.DeviceArray(prob.phi_bytes), L.DeviceArray(prob.phi_bytes)
    y_prev = L.DeviceArray(pred.y_bytes)
    prob.set_tol(TOL_SOLVE, warm=False)
    prob.solve()
    prob.ctx.mg_phi_read(phi_prev)
    pred.run()
    y_prev.copy_from(pred.y, pred.y_bytes)
    names = ("zero", "previous", "learned", "previous_plus_learned_change") + \
        (("previous_plus_learned_delta",) if pdel else ())
    cyc = {k: [] for k in names}
    sec = {k: [] for k in names}
    t_inf = []
    dms = []
    m_last = prob.state()
    for _ in range(states):
        for _ in range(stride):
            prob.relax_step(p_llg)
        m_now = prob.state()
        dms.append(float(np.abs(m_now - m_last).max()))
        s_delta = set_delta(prob, m_now, m_last) if pdel else 0.0
        m_last = m_now
        prob.set_tol(TOL_SOLVE, warm=False)
        t0 = time.perf_counter()
        cyc["zero"].append(prob.solve())
        sec["zero"].append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        pred.run()
        t_inf.append(time.perf_counter() - t0)
        prob.set_tol(TOL_SOLVE, warm=True)
        for name in names[1:]:
            if name == "previous":
                guess.copy_from(phi_prev, prob.phi_bytes)
                extra = 0.0
            elif name == "learned":
                t0 = time.perf_counter()
                pred.potential(pred.y, guess)
                extra = t_inf[-1] + time.perf_counter() - t0
            elif name == "previous_plus_learned_delta":
                t0 = time.perf_counter()
                pdel.run(prob.dm)
                guess.copy_from(phi_prev, prob.phi_bytes)
                pdel.potential(pdel.y, guess, s_delta, True)
                extra = time.perf_counter() - t0
            else:
                t0 = time.perf_counter()
                guess.copy_from(phi_prev, prob.phi_bytes)
                pred.potential(pred.y, guess, 1.0, True)
                pred.potential(y_prev, guess, -1.0, True)
                extra = t_inf[-1] + time.perf_counter() - t0
            prob.ctx.mg_phi_write(guess)
            t0 = time.perf_counter()
            cyc[name].append(prob.solve())
            sec[name].append(time.perf_counter() - t0 + extra)
        prob.ctx.mg_phi_read(phi_prev)        # converged potential of this state
        y_prev.copy_from(pred.y, pred.y_bytes)
    out = {"workload": f"{prob.nx}x{prob.ny} film, 3 nm cells, holes pitch {pitch * 1e9:.0f} nm r 24 nm, "
                       f"box {prob.px}x{prob.py}x{prob.pz}, tol_rel {TOL_SOLVE:g} (>= 1 V-cycle), "
                       f"{states} relaxation states {stride} step(s) apart after {skip} steps, unseen seed",
           "inference_ms": round(1e3 * float(np.mean(t_inf)), 3),
           "max_dm_between_states_mean": round(float(np.mean(dms)), 5), "dt_end": p_llg.dt}
    for k in names:
        out[k] = {"v_cycles_mean": round(float(np.mean(cyc[k])), 3), "v_cycles": cyc[k],
                  "ms_mean_incl_inference": round(1e3 * float(np.mean(sec[k])), 3)}
    prev = out["previous"]["v_cycles_mean"]
    out["learned_vs_previous_cycles"] = round(out["learned"]["v_cycles_mean"] / prev, 3) if prev else None
    out["learned_change_vs_previous_cycles"] = (
        round(out["previous_plus_learned_change"]["v_cycles_mean"] / prev, 3) if prev else None)
    out["learned_vs_zero_cycles"] = round(out["learned"]["v_cycles_mean"] / out["zero"]["v_cycles_mean"], 3)
    if pdel:
        out["learned_delta_vs_previous_cycles"] = (
            round(out["previous_plus_learned_delta"]["v_cycles_mean"] / prev, 3) if prev else None)
        pdel.net.close()
    pred.net.close()
    return out



def _smooth121(c: np.ndarray) -> np.ndarray:
    """[1 2 1] / 4 along the last two axes (y, x), edges clamped."""
    for ax in (1, 2):
        lo = np.concatenate([np.take(c, [0], axis=ax), np.take(c, range(c.shape[ax] - 1), axis=ax)], axis=ax)
        hi = np.concatenate([np.take(c, range(1, c.shape[ax]), axis=ax), np.take(c, [-1], axis=ax)], axis=ax)
        c = 0.25 * lo + 0.5 * c + 0.25 * hi
    return c


def guess_error_study(prob: Problem, params_delta, states: int = 10, seed: int = 424242, pitch: float = 96e-9,
                      skip: int = 5, stride: int = 1, eps=(0.0, 0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 0.7, 1.0),
                      tol_exact: float = 1e-10, smooth=(2,)):
    """How accurate an initial guess must be to save V-cycles, against what
    the trained change network reaches, along the relaxation of compare().

    For each state, phi* is its potential converged to tol_exact (from the
    previous potential) and phi_prev the previous state's potential at
    tol_rel 1e-6, as in compare().  The guesses
        g_eps = phi_prev + (1 - eps) (phi* - phi_prev)
    have error eps times the previous-step guess's error (eps = 1 is the
    previous-step guess, eps = 0 the exact potential); the learned guess is
    phi_prev + s N((m - m_prev) / s) and its error ratio
    ||g - phi*|| / ||phi_prev - phi*|| (L2 over the box) is measured.  Every
    guess is iterated to tol_rel 1e-6 (at least one V-cycle)."""
    pdel = Predictor(prob, params_delta)
    prob.film(seed, pitch)
    p_llg = L.LLGParams(alpha=0.5, dt=1e-13)
    for _ in range(skip):
        prob.relax_step(p_llg)
    nphi = prob.phi_bytes // 8
    phi_prev, guess = L.DeviceArray(prob.phi_bytes), L.DeviceArray(prob.phi_bytes)
    prob.set_tol(TOL_SOLVE, warm=False)
    prob.solve()
    prob.ctx.mg_phi_read(phi_prev)
    m_last = prob.state()
    cyc = {e: [] for e in eps}
    cyc_net, ratio_net, ratio_inf, worst = [], [], [], []
    cyc_s = {k: [] for k in smooth}
    ratio_s = {k: [] for k in smooth}
    for _ in range(states):
        for _ in range(stride):
            prob.relax_step(p_llg)
        m_now = prob.state()
        s_delta = set_delta(prob, m_now, m_last)
        m_last = m_now
        h_prev = phi_prev.download(np.float64, nphi)
        prob.ctx.mg_phi_write(phi_prev)
        prob.set_tol(tol_exact, warm=True)
        prob.solve()
        prob.ctx.mg_phi_read(guess)
        h_star = guess.download(np.float64, nphi)
        e_prev = float(np.linalg.norm(h_prev - h_star))
        prob.set_tol(TOL_SOLVE, warm=True)
        for e in eps:
            guess.upload(h_prev + (1.0 - e) * (h_star - h_prev))
            prob.ctx.mg_phi_write(guess)
            cyc[e].append(prob.solve())
        pdel.run(prob.dm)
        guess.copy_from(phi_prev, prob.phi_bytes)
        pdel.potential(pdel.y, guess, s_delta, True)
        g = guess.download(np.float64, nphi)
        ratio_net.append(float(np.linalg.norm(g - h_star)) / e_prev if e_prev > 0 else 0.0)
        einf_prev = float(np.abs(h_prev - h_star).max())
        err_net = np.abs(g - h_star)
        ratio_inf.append(float(err_net.max()) / einf_prev if einf_prev > 0 else 0.0)
        kz, jy, ix = np.unravel_index(int(err_net.argmax()), (prob.pz, prob.py, prob.px))
        worst.append((int(kz - prob.offz), int(jy - prob.offy), int(ix - prob.offx)))
        prob.ctx.mg_phi_write(guess)
        cyc_net.append(prob.solve())
        # the learned correction low-pass filtered ([1 2 1]/4 along x and y, k times)
        corr = (g - h_prev).reshape(prob.pz, prob.py, prob.px)
        for k in smooth:
            c = corr
            for _ in range(k):
                c = _smooth121(c)
            gs = h_prev + c.reshape(-1)
            ratio_s[k].append(float(np.linalg.norm(gs - h_star)) / e_prev if e_prev > 0 else 0.0)
            guess.upload(gs)
            prob.ctx.mg_phi_write(guess)
            cyc_s[k].append(prob.solve())
        # the previous-step guess of the next state: this state at tol_rel 1e-6
        prob.ctx.mg_phi_write(phi_prev)
        prob.solve()
        prob.ctx.mg_phi_read(phi_prev)
    pdel.net.close()
    prev = float(np.mean(cyc[1.0]))
    return {"workload": f"{prob.nx}x{prob.ny} film, holes pitch {pitch * 1e9:.0f} nm, tol_rel {TOL_SOLVE:g}, "
                        f"{states} states {stride} step(s) apart, exact potential at tol_rel {tol_exact:g}",
            "v_cycles_by_error_ratio": {f"{e:g}": round(float(np.mean(cyc[e])), 2) for e in eps},
            "learned_change": {"error_ratio_mean": round(float(np.mean(ratio_net)), 4),
                               "error_ratio_min": round(float(np.min(ratio_net)), 4),
                               "max_norm_error_ratio_mean": round(float(np.mean(ratio_inf)), 4),
                               "worst_cell_vs_film_origin_kji": worst[:4],
                               "v_cycles": round(float(np.mean(cyc_net)), 2)},
            "learned_change_smoothed": {f"{k}": {"error_ratio_mean": round(float(np.mean(ratio_s[k])), 4),
                                                 "v_cycles": round(float(np.mean(cyc_s[k])), 2)} for k in smooth},
            "previous_step_v_cycles": round(prev, 2),
            "v_cycles_for_30pct_fewer": round(0.7 * prev, 2)}

def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--steps", type=int, default=100, help="relaxation steps (pairs) per run")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None, help="JSON file for the training log and comparison")
    ap.add_argument("--delta", action="store_true",
                    help="train on state differences (phi of m - m_prev); compares with the state network "
                         "at --abs-weights")
    ap.add_argument("--abs-weights", default=DEFAULT_WEIGHTS)
    a = ap.parse_args()
    prob = Problem()
    t0 = time.perf_counter()
    data = record(prob, a.runs, a.steps, seed0=1 if a.delta else 0, delta=a.delta)
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, losses = train(prob, data, a.train_steps, a.batch, a.lr)
    t_train = time.perf_counter() - t0
    params = net.get_params()
    net.close()
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.save(a.out, params)
    n = data.n
    data.x.close()
    data.t.close()
    if a.delta:
        p_abs = np.load(a.abs_weights)
        cmp_ = {f"stride_{k}": compare(prob, p_abs, states=20, stride=k, params_delta=params) for k in (1, 20)}
    else:
        cmp_ = {f"stride_{k}": compare(prob, params, states=20, stride=k) for k in (1, 20)}
    rep = {"pairs": n, "runs": a.runs, "steps_per_run": a.steps, "record_s": round(t_rec, 1),
           "train_steps": a.train_steps, "batch": a.batch, "train_s": round(t_train, 1),
           "loss_first_100": float(np.mean(losses[:100])), "loss_last_100": float(np.mean(losses[-100:])),
           "compare": cmp_}
    print(json.dumps(rep))
    if a.report:
        with open(a.report, "w") as f:
            json.dump({**rep, "losses": [float(v) for v in losses]}, f)


if __name__ == "__main__":
    main()
