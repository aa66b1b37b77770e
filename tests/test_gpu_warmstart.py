"""GPU tests of the learned multigrid initial guess (llgmi_ws_*, csrc/unet.hip).

* Forward pass and every parameter gradient against the numpy oracle
  (oracle/unet_oracle.py, bf16 rounding emulated at the device's storage
  points; the oracle's own gradient is pinned by finite differences).  The
  device accumulates in FP32 (the oracle in float64), so a bf16 rounding can
  land on the other side in a few elements: the bars are relative L2 norms,
  forward 1e-5 (measured 6e-8), each layer's gradient 1e-3 (worst measured
  7e-4), all gradients 3e-4 (measured 1.05e-4).  The spatial-tile
  convolutions (the default) sum 32 channels per matrix-core instruction;
  the pixel-row kernel (LLGMI_UNET_TILE=0, 16 per instruction) lands closer
  to the emulation's roundings (4e-5 worst layer, 3.6e-6 in all), and
  against float64 the two are equally accurate (the large-plane test).
* Adam training on the device reduces the loss of a learnable map.
* The trained network's potential, written into the multigrid solver as its
  initial phi, reaches the same residual in fewer V-cycles than the zero
  guess on states it was not trained on.
"""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _setup(L, s: "U.Shape", seed: int):
    from conftest import DeviceBuffer
    net = L.WarmStartNet(s.batch, s.height, s.width, s.out_channels, s.base,
                         valid_width=s.valid_width or s.width, valid_height=s.valid_height or s.height,
                         seed=seed)
    assert net.n_params == U.param_count(s)
    rng = np.random.default_rng(seed)
    p32 = U.init_params(s, seed).astype(np.float32)
    net.set_params(p32)
    assert np.array_equal(net.get_params(), p32)
    x = np.zeros((s.batch, s.height, s.width, 16))
    x[..., :4] = rng.standard_normal((s.batch, s.height, s.width, 4))
    xb = U.bf16_bits(x)
    t = np.zeros((s.batch, s.height, s.width, s.c_pad))
    t[..., :s.out_channels] = rng.standard_normal((s.batch, s.height, s.width, s.out_channels))
    tb = U.bf16_bits(t)
    xd, td = DeviceBuffer(xb.nbytes), DeviceBuffer(tb.nbytes)
    xd.upload(xb)
    td.upload(tb)
    return net, p32.astype(np.float64), U.bf16_from_bits(xb), U.bf16_from_bits(tb), xd, td


SHAPE = U.Shape(batch=2, height=16, width=24, out_channels=20, base=16, valid_width=22, valid_height=14)


@pytest.mark.parametrize("tile", ["1", "0"])
def test_forward_matches_oracle(gpu_lib, tile, monkeypatch):
    from conftest import DeviceBuffer
    monkeypatch.setenv("LLGMI_UNET_TILE", tile)
    s = SHAPE
    net, p, x, _, xd, _ = _setup(gpu_lib, s, 7)
    n = s.batch * s.height * s.width * s.c_pad
    yd = DeviceBuffer(4 * n)
    net.forward(xd, yd)
    y = yd.download(np.float32, n).reshape(s.batch, s.height, s.width, s.c_pad).astype(np.float64)
    ref, _ = U.forward(s, p, x, bf16=True)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"forward relative L2 vs oracle: {rel:.2e}")
    assert rel <= 1e-5, rel
    assert np.all(np.isfinite(y))


@pytest.mark.parametrize("tile", ["1", "0"])
def test_gradients_match_oracle(gpu_lib, tile, monkeypatch):
    monkeypatch.setenv("LLGMI_UNET_TILE", tile)
    s = SHAPE
    net, p, x, t, xd, td = _setup(gpu_lib, s, 8)
    loss = net.loss_grad(xd, td)
    g = net.get_grads().astype(np.float64)
    loss_o, g_o, _ = U.loss_and_grad(s, p, x, t, bf16=True)
    assert abs(loss - loss_o) <= 1e-3 * loss_o, (loss, loss_o)
    o = 0
    worst = 0.0
    for li, (ci, co, _) in enumerate(U.layers(s)):
        k = 9 * ci * co + co
        a, b = g[o:o + k], g_o[o:o + k]
        if np.linalg.norm(b) > 0:
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            worst = max(worst, rel)
            assert rel <= 1e-3, (li, rel)
        else:
            assert np.all(a == 0)
        o += k
    tot = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    print(f"gradient relative L2 vs oracle: total {tot:.2e}, worst layer {worst:.2e}; loss {loss:.6e} vs {loss_o:.6e}")
    assert tot <= 3e-4


def test_training_reduces_loss(gpu_lib):
    """Fit a fixed linear map of the input (a smoothed combination of the
    channels) on one batch: 150 Adam steps cut the loss by 10x."""
    from conftest import DeviceBuffer
    L = gpu_lib
    s = U.Shape(batch=2, height=32, width=32, out_channels=6, base=16)
    net = L.WarmStartNet(s.batch, s.height, s.width, s.out_channels, s.base, seed=3)
    rng = np.random.default_rng(5)
    x = np.zeros((s.batch, s.height, s.width, 16))
    x[..., :4] = rng.standard_normal((s.batch, s.height, s.width, 4))
    mix = rng.standard_normal((4, s.out_channels)) * 0.5
    t = np.zeros((s.batch, s.height, s.width, s.c_pad))
    sm = x[..., :4].copy()
    for _ in range(2):
        sm = 0.2 * (sm + np.roll(sm, 1, 1) + np.roll(sm, -1, 1) + np.roll(sm, 1, 2) + np.roll(sm, -1, 2))
    t[..., :s.out_channels] = sm @ mix
    xb, tb = U.bf16_bits(x), U.bf16_bits(t)
    xd, td = DeviceBuffer(xb.nbytes), DeviceBuffer(tb.nbytes)
    xd.upload(xb)
    td.upload(tb)
    losses = [net.train_step(xd, td, 2e-3) for _ in range(150)]
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])
    with pytest.raises(L.LlgmiError):
        net.train_step(xd, td, -1.0)


def _film(nx, ny, seed, pitch_cells, radius_cells):
    from initial_states import smoothed_random
    m = smoothed_random(nx, ny, seed).reshape(ny, nx, 3).copy()
    yy, xx = np.mgrid[0:ny, 0:nx]
    cx = (xx + 0.5) % pitch_cells - pitch_cells / 2
    cy = (yy + 0.5) % pitch_cells - pitch_cells / 2
    m[cx * cx + cy * cy < radius_cells ** 2] = 0.0
    return m.reshape(-1, 3)


def test_learned_guess_saves_v_cycles(gpu_lib):
    from conftest import DeviceBuffer
    L = gpu_lib
    nx = ny = 64
    cell = (3e-9, 3e-9, 3e-9)
    ms = 8.6e5
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    mat = L.Material(ms=ms, demag=True, demag_method=L.DemagMethod.PoissonMG)
    base = dict(stencil=0, prolong=0, coarse_op=0, smoother=1, bc=2, v_cycles=1, v_cycles_max=300)
    ctx.set_mg_config(L.mg_config(warm_start=0, tol_rel=1e-9, **base))
    ctx.compute_demag_field_poisson_mg(_film(nx, ny, 1, 16, 4), mat)
    px, py, pz, offx, offy, offz = ctx.mg_box()
    H = W = (max(px, py) + 7) // 8 * 8
    scale = ms * cell[2]
    B = 4
    net = L.WarmStartNet(B, H, W, pz, 32, valid_width=px, valid_height=py, seed=11)
    n_train = 48
    xs, ts = DeviceBuffer(n_train * H * W * 16 * 2), DeviceBuffer(n_train * H * W * net.c_pad * 2)
    mdev = DeviceBuffer(3 * nx * ny * 8)
    phi = DeviceBuffer(px * py * pz * 8)
    xb, tb = DeviceBuffer(B * H * W * 16 * 2), DeviceBuffer(B * H * W * net.c_pad * 2)
    sx, st = H * W * 16 * 2, H * W * net.c_pad * 2
    for k in range(n_train):   # converged potentials of random films with holes (pitch 12..20 cells)
        m = _film(nx, ny, 100 + k, 12 + k % 9, 3 + k % 3)
        ctx.compute_demag_field_poisson_mg(m, mat)
        ctx.mg_phi_read(phi)
        ctx.dev_upload(m, mdev)
        net.pack_input(0, mdev, False, nx, ny, offx, offy, xb)
        net.pack_phi(0, phi, px, py, pz, scale, tb)
        xs.copy_from(xb, sx, 0, k * sx)
        ts.copy_from(tb, st, 0, k * st)
    rng = np.random.default_rng(0)
    losses = []
    for it in range(400):
        idx = rng.choice(n_train, B, replace=False)
        for j, k in enumerate(idx):
            xb.copy_from(xs, sx, k * sx, j * sx)
            tb.copy_from(ts, st, k * st, j * st)
        losses.append(net.train_step(xb, tb, 1e-3 if it < 300 else 3e-4))
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
    y = DeviceBuffer(B * H * W * net.c_pad * 4)
    cyc_zero, cyc_net = [], []
    for k in range(6):            # held-out films
        m = _film(nx, ny, 900 + k, 13 + k, 4)
        ctx.set_mg_config(L.mg_config(warm_start=0, tol_rel=1e-6, **base))
        ctx.compute_demag_field_poisson_mg(m, mat)
        cyc_zero.append(ctx.mg_last_stats()["cycles"])
        ctx.dev_upload(m, mdev)
        net.pack_input(0, mdev, False, nx, ny, offx, offy, xb)
        net.forward(xb, y)
        net.unpack_phi(0, y, px, py, pz, scale, phi)
        ctx.set_mg_config(L.mg_config(warm_start=1, tol_rel=1e-6, **base))
        ctx.mg_phi_write(phi)
        ctx.compute_demag_field_poisson_mg(m, mat)
        cyc_net.append(ctx.mg_last_stats()["cycles"])
    print(f"V-cycles to 1e-6: zero guess {cyc_zero}, learned guess {cyc_net}; loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert np.mean(cyc_net) < np.mean(cyc_zero)


def test_workflow_record_train_compare_small(gpu_lib):
    """The recording / training / comparison workflow of warmstart.py on a
    128 x 128 film (runs end to end; every guess reaches the tolerance)."""
    import warmstart as WS
    prob = WS.Problem(128, 128)
    assert prob.H % 8 == 0 and prob.H >= prob.px
    data = WS.record(prob, runs=2, steps=6, log=lambda *_: None)
    t0 = data.t.download(np.uint16, data.st // 2)
    assert data.n == 12 and np.any(t0 != 0)
    net, losses = WS.train(prob, data, steps=60, batch=4, log=lambda *_: None)
    assert np.all(np.isfinite(losses)) and np.mean(losses[-10:]) < np.mean(losses[:10])
    out = WS.compare(prob, net.get_params(), states=3, skip=1)
    for k in ("zero", "previous", "learned", "previous_plus_learned_change"):
        assert len(out[k]["v_cycles"]) == 3 and min(out[k]["v_cycles"]) >= 1
        assert max(out[k]["v_cycles"]) < 300


def test_mask_cells_and_unpack_accumulate(gpu_lib):
    L = gpu_lib
    nx, ny = 16, 8
    ctx = L.Context(L.Grid2D(nx, ny, 1e-9, 1e-9, 1e-9), precision=L.F64)
    rng = np.random.default_rng(2)
    m = rng.standard_normal((nx * ny, 3))
    keep = (rng.random(nx * ny) > 0.3).astype(np.uint8)
    md, kd = L.DeviceArray(m.nbytes), L.DeviceArray(keep.nbytes)
    ctx.dev_upload(m, md)
    kd.upload(keep)
    ctx.dev_mask_cells(md, kd)
    out = ctx.dev_download(md)
    assert np.array_equal(out[keep == 1], m[keep == 1]) and np.all(out[keep == 0] == 0.0)
    net = L.WarmStartNet(1, 16, 24, 5, 16)
    y = rng.standard_normal((16, 24, net.c_pad)).astype(np.float32)
    yd = L.DeviceArray(y.nbytes)
    yd.upload(y)
    px, py, pz = 20, 14, 5
    phi = L.DeviceArray(px * py * pz * 8)
    net.unpack_phi(0, yd, px, py, pz, 2.0, phi)
    ref = 2.0 * y[:py, :px, :pz].astype(np.float64).transpose(2, 0, 1)
    assert np.array_equal(phi.download(np.float64, px * py * pz).reshape(pz, py, px), ref)
    net.unpack_phi(0, yd, px, py, pz, -0.5, phi, accumulate=True)
    assert np.allclose(phi.download(np.float64, px * py * pz).reshape(pz, py, px), 0.75 * ref, rtol=1e-15)
    with pytest.raises(L.LlgmiError):
        net.unpack_phi(0, yd, px, py, 6, 1.0, phi)         # more planes than output channels
    with pytest.raises(L.LlgmiError):
        ctx.dev_mask_cells(md, L.DeviceArray(4))             # keep buffer too short


@pytest.mark.parametrize("h,w", [(240, 264), (256, 264)])
def test_large_planes_as_accurate_as_bf16_emulation(gpu_lib, h, w):
    """On planes of 63360 and 67584 pixels (the second runs 64-pixel waves
    at level 0) the device and the oracle's bf16 emulation each round some
    activations differently (FP32 vs float64 sums), and through the deep
    levels one such rounding reaches much of the image: the two differ by
    about the bf16 noise itself (measured 3.7e-3 relative, against 6e-3 for
    either vs float64).  The bar is therefore accuracy against the float64
    oracle: within 10 % of the bf16 emulation's own error, forward and
    gradients (measured: equal to 3 digits)."""
    from conftest import DeviceBuffer
    s = U.Shape(batch=1, height=h, width=w, out_channels=5, base=16, valid_width=w - 4, valid_height=h - 6)
    net, p, x, t, xd, td = _setup(gpu_lib, s, 21)
    n = s.batch * s.height * s.width * s.c_pad
    yd = DeviceBuffer(4 * n)
    net.forward(xd, yd)
    y = yd.download(np.float32, n).reshape(s.batch, s.height, s.width, s.c_pad).astype(np.float64)
    net.loss_grad(xd, td)
    g = net.get_grads().astype(np.float64)
    _, g_b, y_b = U.loss_and_grad(s, p, x, t, bf16=True)
    _, g_f, y_f = U.loss_and_grad(s, p, x, t, bf16=False)

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"{h}x{w}: forward vs f64 {rel(y, y_f):.3e} (bf16 oracle {rel(y_b, y_f):.3e}), "
          f"gradient vs f64 {rel(g, g_f):.3e} (bf16 oracle {rel(g_b, g_f):.3e})")
    assert rel(y, y_f) <= 1.1 * rel(y_b, y_f) and rel(g, g_f) <= 1.1 * rel(g_b, g_f)
    assert rel(y, y_f) < 2e-2 and rel(g, g_f) < 5e-2
