"""GPU parity of the FFT demag (demag_fft_uniform.rs) through the C ABI.

Bars: FP64 context within 1e-12 of the oracle (max |dB| / max |B|); FP32
context within 1e-5 of the FP64 context.  Known answers mirrored from the
reference: demag_single_cell_cube_is_close_to_minus_mu0_over_3
(demag_fft_uniform.rs:1261-1293), demag_uniform_2x2_* (tests/validation.rs:478-615).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

MU0 = 4e-7 * np.pi


def _rand_m(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("shape,cell", [
    ((1, 1), (1.0, 1.0, 1.0)),
    ((2, 2), (1.0, 1.0, 1.0)),
    ((3, 5), (2e-9, 2e-9, 1e-9)),
    ((16, 8), (4e-9, 4e-9, 3e-9)),
    ((37, 23), (3e-9, 2e-9, 1e-9)),
    ((128, 64), (4e-9, 4e-9, 3e-9)),
    ((1, 40), (5e-9, 5e-9, 5e-9)),
    ((50, 1), (2e-9, 5e-9, 1e-9)),
])
def test_demag_fp64_matches_oracle(gpu_lib, shape, cell):
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = shape
    m = _rand_m(nx * ny, nx * 1000 + ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = np.zeros_like(m)
    ctx.add_demag_field_pbc(m, b, L.Material(ms=8e5, demag=True), 0, 0)
    assert rel_err(b, b_ref) <= 1e-12


def test_demag_accumulates(gpu_lib):
    L = gpu_lib
    nx, ny = 8, 6
    m = _rand_m(nx * ny, 3)
    ctx = L.Context(L.Grid2D(nx, ny, 2e-9, 2e-9, 1e-9), precision=L.F64)
    mat = L.Material(ms=8e5, demag=True)
    b0 = ctx.compute_demag_field(m, mat)
    b = np.full_like(m, 0.25)
    ctx.add_demag_field(m, b, mat)
    np.testing.assert_allclose(b - 0.25, b0, rtol=0, atol=1e-15)
    off = np.full_like(m, 0.5)
    ctx.add_demag_field(m, off, L.Material(ms=8e5, demag=False))
    assert (off == 0.5).all()


def test_demag_single_cell_cube_is_close_to_minus_mu0_over_3(gpu_lib):
    L = gpu_lib
    ctx = L.Context(L.Grid2D(1, 1, 1.0, 1.0, 1.0), precision=L.F64)
    b = np.zeros((1, 3))
    ctx.add_demag_field(np.array([[0.0, 0.0, 1.0]]), b, L.Material(ms=1.0, demag=True))
    assert abs(b[0, 2] - (-MU0 / 3.0)) < 1e-3


@pytest.mark.parametrize("prec", [1, 0])
def test_demag_uniform_2x2_symmetry_and_factors(gpu_lib, prec):
    L = gpu_lib
    ctx = L.Context(L.Grid2D(2, 2, 1.0, 1.0, 1.0), precision=prec)
    mat = L.Material(ms=1.0, demag=True)
    n = []
    for comp in range(3):
        m = np.zeros((4, 3))
        m[:, comp] = 1.0
        b = ctx.compute_demag_field(m, mat)
        others = [c for c in range(3) if c != comp]
        if comp == 2:
            assert np.abs(b[:, :2]).max() < 1e-10 * MU0 * 10
        assert np.ptp(b[:, comp]) < 1e-10 * MU0 * 10
        n.append(-b[0, comp] / MU0)
        del others
    nxx, nyy, nzz = n
    assert 0 < nxx < 1.2 and 0 < nyy < 1.2 and 0.05 < nzz < 1.05
    assert 0.5 < nxx + nyy + nzz < 2.5


@pytest.mark.parametrize("shape,cell", [
    ((128, 64), (4e-9, 4e-9, 3e-9)),
    ((200, 96), (2e-9, 2e-9, 1e-9)),
    ((256, 256), (3e-9, 3e-9, 3e-9)),
])
def test_full_field_fp32_vs_fp64(gpu_lib, shape, cell):
    """FP32 context against the FP64 context on the same m (per term and total)."""
    L = gpu_lib
    nx, ny = shape
    m = _rand_m(nx * ny, 99)
    # smooth the state a little so exchange does not dominate everything
    mat = L.Material(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, dmi=3e-3, demag=True)
    p = L.LLGParams(b_ext=(-24.6e-3, 4.3e-3, 0.0))
    c64 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    c32 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    b64 = c64.build_h_eff(m, p, mat)
    b32 = c32.build_h_eff(m, p, mat)
    assert rel_err(b32, b64) <= 1e-5
    d64 = c64.compute_demag_field(m, mat)
    d32 = c32.compute_demag_field(m, mat)
    assert rel_err(d32, d64) <= 1e-5


@pytest.mark.parametrize("radix", ["8", "16"])
@pytest.mark.parametrize("shape,cell", [
    ((128, 64), (4e-9, 4e-9, 3e-9)),
    ((300, 3), (5e-9, 5e-9, 5e-9)),
    ((129, 130), (3e-9, 2e-9, 1e-9)),
])
def test_demag_forced_radix_fp64_matches_oracle(gpu_lib, monkeypatch, radix, shape, cell):
    """Both first/last-stage radices of the FFT passes (LLGMI_FFT_RADIX)."""
    from oracle import oracle as O
    monkeypatch.setenv("LLGMI_FFT_RADIX", radix)
    L = gpu_lib
    nx, ny = shape
    m = _rand_m(nx * ny, 7 * nx + ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-12



# Long FFT lengths: padded 4096 and 8192 along each axis (rows are the x
# transforms, columns the y transforms), up to 8192 in both precisions.  FP64
# columns of 8192 (ny > 2048) take the split column pass (one sequence per
# block, forward spectra through HBM).
@pytest.mark.parametrize("radix", ["auto", "8", "16"])
@pytest.mark.parametrize("shape", [(8, 1536), (1536, 8), (2100, 8), (5, 1100), (8, 2100), (3, 4096)])
def test_demag_long_lengths_fp64_matches_oracle(gpu_lib, monkeypatch, radix, shape):
    from oracle import oracle as O
    if radix != "auto":
        monkeypatch.setenv("LLGMI_FFT_RADIX", radix)
    L = gpu_lib
    nx, ny = shape
    cell = (3e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 11 * nx + ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-12


def test_demag_fp64_1536_square_matches_oracle(gpu_lib):
    """1536 x 1536 pads to 4096 x 4096 (both passes at length 4096)."""
    from oracle import oracle as O
    L = gpu_lib
    nx = ny = 1536
    cell = (2e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 1536)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-12


@pytest.mark.parametrize("shape", [(8, 2100), (4096, 4096)])
def test_demag_fp32_length_8192_matches_oracle(gpu_lib, shape):
    """FP32 columns (and rows) of padded length 8192, against the FP64 oracle
    at the FP32 bar."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = shape
    cell = (3e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, nx + 3 * ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-5


# Lengths 3 * 2^k: a grid whose 2n is 3 * 2^k (or closer to it than to the
# next power of two) pads to 3 * 2^k, as the reference's 2n does; three
# register transforms of 2^k points and a radix-3 combination.
@pytest.mark.parametrize("radix", ["auto", "8", "16"])
@pytest.mark.parametrize("shape", [(1536, 8), (8, 1536), (384, 384), (1400, 5), (5, 1400), (200, 700)])
def test_demag_radix3_lengths_fp64_match_oracle(gpu_lib, monkeypatch, radix, shape):
    from oracle import oracle as O
    monkeypatch.setenv("LLGMI_FFT_PAD", "3")     # 3 * 2^k whenever shorter, also below 3072
    if radix != "auto":
        monkeypatch.setenv("LLGMI_FFT_RADIX", radix)
    L = gpu_lib
    nx, ny = shape
    cell = (3e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 13 * nx + ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-12


def test_demag_fp64_1536_square_pads_to_3072(gpu_lib, monkeypatch):
    """1536 x 1536 (2n = 3072 = 3 * 2^10 along both axes) against the oracle,
    and against the same grid padded to powers of two (LLGMI_FFT_PAD=pow2)."""
    from oracle import oracle as O
    L = gpu_lib
    nx = ny = 1536
    cell = (2e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 3072)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    mat = L.Material(ms=8e5, demag=True)
    b3 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64).compute_demag_field(m, mat)
    monkeypatch.setenv("LLGMI_FFT_PAD", "pow2")
    b2 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64).compute_demag_field(m, mat)
    assert rel_err(b3, b_ref) <= 1e-12
    assert rel_err(b3, b2) <= 1e-12


@pytest.mark.parametrize("shape", [(3000, 8), (8, 3000), (768, 768)])
def test_demag_radix3_lengths_fp32_match_oracle(gpu_lib, monkeypatch, shape):
    """FP32 at padded lengths 6144 (rows and columns) and 1536."""
    from oracle import oracle as O
    monkeypatch.setenv("LLGMI_FFT_PAD", "3")
    L = gpu_lib
    nx, ny = shape
    cell = (3e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, nx + 7 * ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-5


@pytest.mark.parametrize("prec", [1, 0])
def test_full_field_radix3_lengths_fused_and_unfused(gpu_lib, monkeypatch, prec):
    """The whole field (all terms, FFT demag padded to 768 x 1536) with the
    fused and the unfused inverse row pass, against the oracle."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = 384, 768
    cell = (2e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 77)
    okw = dict(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, dmi=3e-3, demag=True)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = oc.build_h_eff(m, O.Params(alpha=0.3, b_ext=(-24.6e-3, 4.3e-3, 0.0)), O.Material(**okw))
    monkeypatch.setenv("LLGMI_FFT_PAD", "3")
    for fused in ("0", "1"):
        monkeypatch.setenv("LLGMI_FFT_FUSED", fused)
        ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=prec)
        b = ctx.build_h_eff(m, L.LLGParams(b_ext=(-24.6e-3, 4.3e-3, 0.0)), L.Material(**okw))
        assert rel_err(b, b_ref) <= (1e-12 if prec == 1 else 1e-5), fused


def test_demag_fp64_4096_square_matches_oracle(gpu_lib):
    """4096 x 4096 in FP64: rows of 8192 and split columns of 8192."""
    from oracle import oracle as O
    L = gpu_lib
    nx = ny = 4096
    cell = (3e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 4096)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag(m, b_ref, O.Material(ms=8e5, demag=True))
    del oc
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = ctx.compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, b_ref) <= 1e-12


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("shape", [(8, 4097), (4097, 8)])
def test_demag_beyond_built_lengths_is_unsupported(gpu_lib, prec, shape):
    """A padded length of 16384 (nx or ny > 4096) is refused with
    LLGMI_ERR_UNSUPPORTED, before any launch."""
    L = gpu_lib
    nx, ny = shape
    ctx = L.Context(L.Grid2D(nx, ny, 3e-9, 2e-9, 1e-9), precision=prec)
    with pytest.raises(L.LlgmiError) as e:
        ctx.compute_demag_field(_rand_m(nx * ny, 5), L.Material(ms=8e5, demag=True))
    assert e.value.code == L.ERR_UNSUPPORTED

@pytest.mark.parametrize("prec", [1, 0])
def test_demag_radix_variants_agree_1024(gpu_lib, monkeypatch, prec):
    """At the 1024 x 1024 workload the radix-8 and radix-16 passes agree to
    rounding, and the FP32 result stays within 1e-5 of FP64."""
    L = gpu_lib
    nx = ny = 1024
    cell = (2e-9, 2e-9, 1e-9)
    m = _rand_m(nx * ny, 11)
    mat = L.Material(ms=8e5, demag=True)
    out = {}
    for radix in ("8", "16"):
        monkeypatch.setenv("LLGMI_FFT_RADIX", radix)
        out[radix] = L.Context(L.Grid2D(nx, ny, *cell), precision=prec).compute_demag_field(m, mat)
    tol = 1e-13 if prec == 1 else 2e-6
    assert rel_err(out["8"], out["16"]) <= tol
    if prec == 0:
        monkeypatch.setenv("LLGMI_FFT_RADIX", "16")
        ref = L.Context(L.Grid2D(nx, ny, *cell), precision=1).compute_demag_field(m, mat)
        assert rel_err(out["8"], ref) <= 1e-5


def test_fft_radix_env_rejects_other_values(gpu_lib, monkeypatch):
    L = gpu_lib
    monkeypatch.setenv("LLGMI_FFT_RADIX", "32")
    ctx = L.Context(L.Grid2D(8, 8, 1e-9, 1e-9, 1e-9), precision=L.F64)
    with pytest.raises(Exception, match="LLGMI_FFT_RADIX"):
        ctx.compute_demag_field(np.tile([0.0, 0.0, 1.0], (64, 1)), L.Material(ms=8e5, demag=True))


@pytest.mark.parametrize("prec", [1, 0])
@pytest.mark.parametrize("shape,cell", [
    ((64, 48), (3e-9, 3e-9, 2e-9)),
    ((300, 200), (2e-9, 2e-9, 1e-9)),
    ((512, 512), (2e-9, 2e-9, 1e-9)),
])
def test_fused_inverse_row_and_local_field_match_unfused(gpu_lib, monkeypatch, prec, shape, cell):
    """The inverse row FFT fused with the local field (LLGMI_FFT_FUSED=1)
    against the unfused pair (=0): field builds with and without a geometry
    mask, and RK45 steps (torque variant, FP64).  The two compile the same
    arithmetic into different kernels, where the compiler may contract
    multiply-adds differently, so they agree to rounding: 1e-13 (FP64) /
    2e-6 (FP32) of max |B|."""
    L = gpu_lib
    nx, ny = shape
    m = _rand_m(nx * ny, 21)
    geom = np.ones(nx * ny, dtype=bool)
    geom[::7] = False
    mat = L.Material(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, dmi=3e-3, demag=True)
    p = L.LLGParams(alpha=0.3, dt=2e-14, b_ext=(-24.6e-3, 4.3e-3, 0.0))
    tol = 1e-13 if prec == 1 else 2e-6
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("LLGMI_FFT_FUSED", fused)
        ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=prec)
        b1 = ctx.build_h_eff(m, p, mat)
        b2 = ctx.build_h_eff_masked_geom(m, p, mat, L.FieldMask.Full, geom)
        mm = m.copy()
        pp = L.LLGParams(alpha=0.3, dt=2e-14, b_ext=(-24.6e-3, 4.3e-3, 0.0))
        steps = [ctx.step_llg_rk45_recompute_field_masked_adaptive(mm, pp, mat, L.FieldMask.Full, 1e-5, 0.8,
                                                                   1e-18, 1e-12) for _ in range(3)]
        out[fused] = (b1, b2, mm, steps)
        ctx.close()
    assert rel_err(out["1"][0], out["0"][0]) <= tol
    assert rel_err(out["1"][1], out["0"][1]) <= tol
    if prec == 0:
        # FP32 steps on this rough random state are not compared: the
        # embedded error estimate is a small difference of large slopes, so
        # FP32 rounding moves eps enough to change the adaptive step size
        return
    assert np.abs(out["1"][2] - out["0"][2]).max() <= 1e-12
    for (e1, a1, d1), (e0, a0, d0) in zip(out["1"][3], out["0"][3]):
        assert a1 == a0
        assert abs(e1 - e0) <= 1e-6 * max(abs(e0), 1e-300) + 1e-15
        assert abs(d1 - d0) <= 1e-6 * abs(d0)


@pytest.mark.parametrize("shape,pbc", [
    ((8, 6), (2, 0)),
    ((16, 8), (0, 3)),
    ((12, 10), (2, 2)),
    ((1, 1), (5, 5)),
    ((33, 17), (1, 0)),
    ((1, 24), (3, 0)),
    ((64, 64), (4, 4)),
])
def test_demag_pbc_fp64_matches_oracle(gpu_lib, shape, pbc):
    """add_demag_field_pbc with periodic images (demag_fft_uniform.rs:941-1005):
    the oracle convolves over exactly n cells on a periodic axis, the GPU with
    a zero-padded FFT and the folded periodic kernel; both are the circular
    convolution of period n."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = shape
    print(f"pbc {shape} {pbc}", flush=True)
    cell = (4e-9, 3e-9, 2e-9)
    m = _rand_m(nx * ny, nx * 31 + ny)
    oc = O.Context(O.Grid(nx, ny, *cell))
    b_ref = np.zeros_like(m)
    oc.add_demag_pbc(m, b_ref, O.Material(ms=8e5, demag=True), *pbc)
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    b = np.zeros_like(m)
    ctx.add_demag_field_pbc(m, b, L.Material(ms=8e5, demag=True), *pbc)
    assert rel_err(b, b_ref) <= 1e-12
    # the open operator of the same context is untouched by the periodic one
    bo = np.zeros_like(m)
    ctx.add_demag_field_pbc(m, bo, L.Material(ms=8e5, demag=True), 0, 0)
    bo_ref = np.zeros_like(m)
    oc.add_demag(m, bo_ref, O.Material(ms=8e5, demag=True))
    assert rel_err(bo, bo_ref) <= 1e-12
    # FP32 against FP64
    c32 = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    b32 = np.zeros_like(m)
    c32.add_demag_field_pbc(m, b32, L.Material(ms=8e5, demag=True), *pbc)
    assert rel_err(b32, b) <= 1e-5


def _write_reference_cache(path, nx, ny, cell, pbc, kspace):
    """An LLGDMAG v2 file laid out as write_kernel_kspace writes it
    (demag_fft_uniform.rs:282-320), holding the oracle's k-space kernels."""
    import struct
    px, py = kspace.shape[2], kspace.shape[1]
    with open(path, "wb") as f:
        f.write(b"LLGDMAG\0")
        f.write(struct.pack("<7I", 2, nx, ny, pbc[0], pbc[1], px, py))
        f.write(struct.pack("<4d", cell[0], cell[1], cell[2], 10.0))
        for c in range(4):
            f.write(np.ascontiguousarray(kspace[c]).astype("<c16").tobytes())


def _read_cache(path):
    import struct
    raw = open(path, "rb").read()
    assert raw[:8] == b"LLGDMAG\0"
    hdr = struct.unpack("<7I", raw[8:36]) + struct.unpack("<4d", raw[36:68])
    px, py = hdr[5], hdr[6]
    k = np.frombuffer(raw[68:], dtype="<c16")
    assert k.size == 4 * px * py
    return hdr, k.reshape(4, py, px)


@pytest.mark.parametrize("shape,pbc", [((16, 8), (0, 0)), ((30, 7), (0, 0)), ((12, 10), (2, 0)), ((9, 6), (1, 2))])
def test_kernel_cache_interchange_with_the_reference_layout(gpu_lib, tmp_path, shape, pbc):
    """try_load_kernel_kspace / write_kernel_kspace (demag_fft_uniform.rs:282-429):
    a cache file at the reference's FFT sizes (2n open, n periodic; not powers of
    two here) written from the oracle's k-space kernels loads into a context and
    gives the oracle's field; the context's own file holds the oracle's k-space
    kernels."""
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = shape
    cell = (4e-9, 3e-9, 2e-9)
    print(f"cache {shape} {pbc}", flush=True)
    od = O.Demag(O.Grid(nx, ny, *cell), *pbc)
    ks = od.kspace()
    src = str(tmp_path / "ref.bin")
    _write_reference_cache(src, nx, ny, cell, pbc, ks)
    m = _rand_m(nx * ny, 5 * nx + ny)
    b_ref = np.zeros_like(m)
    od.add_field(m, b_ref, 8e5)
    ctx = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F64)
    assert ctx.demag_cache_load(*pbc, path=src)
    b = np.zeros_like(m)
    ctx.add_demag_field_pbc(m, b, L.Material(ms=8e5, demag=True), *pbc)
    assert rel_err(b, b_ref) <= 1e-12
    # a fresh context writes the same kernels
    out = str(tmp_path / "sub" / "mine.bin")
    fresh = L.Context(L.Grid2D(nx, ny, *cell), precision=L.F32)
    fresh.demag_cache_save(*pbc, path=out)
    hdr, k = _read_cache(out)
    assert hdr[:7] == (2, nx, ny, pbc[0], pbc[1], ks.shape[2], ks.shape[1])
    assert hdr[7:] == (cell[0], cell[1], cell[2], 10.0)
    assert np.abs(k - ks).max() <= 1e-12 * np.abs(ks).max()
    # a header for another grid is not loaded (nothing changes)
    other = L.Context(L.Grid2D(nx + 1, ny, *cell), precision=L.F64)
    assert not other.demag_cache_load(*pbc, path=src)
    assert not other.demag_cache_load(*pbc, path=str(tmp_path / "absent.bin"))


def test_kernel_cache_automatic_with_env(gpu_lib, tmp_path, monkeypatch):
    """LLGMI_DEMAG_CACHE=1: the operator build reads and writes the reference's
    cache file (out/demag_cache/... in the working directory)."""
    L = gpu_lib
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("LLGMI_DEMAG_CACHE", "1")
    g = L.Grid2D(20, 12, 4e-9, 4e-9, 3e-9)
    m = _rand_m(240, 77)
    a = L.Context(g, precision=L.F64).compute_demag_field(m, L.Material(ms=8e5, demag=True))
    path = tmp_path / L.demag_cache_path(g)
    assert path.exists()
    b = L.Context(g, precision=L.F64).compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, a) <= 1e-13


def test_kernel_cache_automatic_rebuilds_a_truncated_file(gpu_lib, tmp_path, monkeypatch):
    """LLGMI_DEMAG_CACHE=1 with a short cache file at the automatic path: the
    file is a miss, as try_load_kernel_kspace(..).unwrap_or(false) makes it
    (demag_fft_uniform.rs:499-512); the operator is built and the file written
    again, whole."""
    L = gpu_lib
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("LLGMI_DEMAG_CACHE", "1")
    g = L.Grid2D(20, 12, 4e-9, 4e-9, 3e-9)
    m = _rand_m(240, 78)
    a = L.Context(g, precision=L.F64).compute_demag_field(m, L.Material(ms=8e5, demag=True))
    path = tmp_path / L.demag_cache_path(g)
    whole = path.read_bytes()
    path.write_bytes(whole[: len(whole) // 3])          # truncated mid-kernel
    b = L.Context(g, precision=L.F64).compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(b, a) <= 1e-13
    assert path.read_bytes() == whole
    path.write_bytes(whole[:20])                        # truncated mid-header
    c = L.Context(g, precision=L.F64).compute_demag_field(m, L.Material(ms=8e5, demag=True))
    assert rel_err(c, a) <= 1e-13
    assert path.read_bytes() == whole


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_cache_load_between_graph_replayed_steps(gpu_lib, tmp_path, prec, monkeypatch):
    """Loading a kernel cache rebuilds the operator that captured RK45 attempts
    read: steps replayed from graphs, then demag_cache_load, then more steps,
    stay bitwise equal to the same sequence launched eagerly."""
    from conftest import DeviceBuffer
    L = gpu_lib
    P, es = (L.F32, 4) if prec == "f32" else (L.F64, 8)
    nx, ny, cell = 24, 16, (3e-9, 3e-9, 2e-9)
    m0 = _rand_m(nx * ny, 91)
    mat = L.Material(ms=8e5, a_ex=1.3e-11, k_u=2e4, demag=True)
    src = str(tmp_path / "k.bin")
    L.Context(L.Grid2D(nx, ny, *cell), precision=P).demag_cache_save(0, 0, path=src)
    cg = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    ce = L.Context(L.Grid2D(nx, ny, *cell), precision=P)
    bg, be = DeviceBuffer(3 * nx * ny * es), DeviceBuffer(3 * nx * ny * es)
    cg.dev_upload(m0, bg)
    ce.dev_upload(m0, be)
    pg = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    pe = L.LLGParams(alpha=0.1, dt=1e-13, b_ext=(0.0, 0.0, 0.02))
    for step in range(10):
        if step == 5:
            assert cg.demag_cache_load(0, 0, path=src)
            assert ce.demag_cache_load(0, 0, path=src)
        monkeypatch.delenv("LLGMI_RK45_GRAPH", raising=False)
        rg = cg.dev_rk45_step(bg, pg, mat, 2, 1e-5, 0.8, 1e-18, 1e-11)
        monkeypatch.setenv("LLGMI_RK45_GRAPH", "0")
        re_ = ce.dev_rk45_step(be, pe, mat, 2, 1e-5, 0.8, 1e-18, 1e-11)
        assert rg == re_, step
        assert np.array_equal(cg.dev_download(bg), ce.dev_download(be)), step


def test_fft_pad_env_rejects_other_values(gpu_lib, monkeypatch):
    L = gpu_lib
    monkeypatch.setenv("LLGMI_FFT_PAD", "5")
    ctx = L.Context(L.Grid2D(16, 16, 3e-9, 3e-9, 3e-9), precision=L.F64)
    with pytest.raises(L.LlgmiError) as e:
        ctx.compute_demag_field(_rand_m(256, 1), L.Material(ms=8e5, demag=True))
    assert e.value.code == L.ERR_INVALID
