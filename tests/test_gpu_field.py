"""GPU parity of the local field terms (Zeeman, exchange, anisotropy, DMI) and
the field builder masks, through the C ABI, against the CPU oracle.

Bars: FP64 context within 1e-12 (max |dB| / max |B|), FP32 within 1e-5.
Reference tests mirrored: dmi_field_flips_sign_with_d (tests/validation.rs:395-472).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _pair(**kw):
    from oracle import oracle as O
    import llgmi
    return O.Material(**kw), llgmi.Material(**kw)


def _params(**kw):
    from oracle import oracle as O
    import llgmi
    return O.Params(**kw), llgmi.LLGParams(**kw)


def _rand_m(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


MATS = [
    dict(ms=8.0e5, a_ex=1.3e-11),
    dict(ms=5.8e5, a_ex=1.5e-11, k_u=8e5, easy_axis=(0.0, 0.0, 1.0), dmi=3e-3),
    dict(ms=8.6e5, a_ex=1.0e-11, k_u=-2e5, easy_axis=(0.6, 0.0, 0.8), dmi=-1.5e-3),
    dict(ms=8.0e5, a_ex=0.0, k_u=5e5, easy_axis=(0.0, 1.0, 0.0), dmi=2e-3),  # A=0: DMI off
    dict(ms=8.0e5, a_ex=1.3e-11, dmi=0.0),                                    # D=Some(0)
]
SHAPES = [(37, 23), (1, 9), (9, 1), (1, 1), (64, 32), (5, 4)]


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-5)])
@pytest.mark.parametrize("mi", range(len(MATS)))
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mask", [0, 1, 2])
def test_build_h_eff_masked_matches_oracle(gpu_lib, prec, tol, mi, shape, mask):
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = shape
    om, gm = _pair(**MATS[mi])
    op, gp = _params(b_ext=(-24.6e-3, 4.3e-3, 0.01))
    m = _rand_m(nx * ny, 100 + mi)
    oc = O.Context(O.Grid(nx, ny, 2e-9, 3e-9, 1e-9))
    b_ref = oc.build_h_eff(m, op, om, mask)
    ctx = L.Context(L.Grid2D(nx, ny, 2e-9, 3e-9, 1e-9), precision=prec)
    b = ctx.build_h_eff_masked(m, gp, gm, mask)
    assert rel_err(b, b_ref) <= tol, (shape, mi, mask)


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-5)])
@pytest.mark.parametrize("mi", [0, 1, 2])
def test_build_h_eff_masked_geom_matches_oracle(gpu_lib, prec, tol, mi):
    from oracle import oracle as O
    L = gpu_lib
    nx, ny = 41, 29
    rng = np.random.default_rng(7 + mi)
    geom = rng.random(nx * ny) > 0.3
    om, gm = _pair(**MATS[mi])
    op, gp = _params(b_ext=(0.01, -0.02, 0.03))
    m = _rand_m(nx * ny, 11 + mi)
    m[~geom] = 0.0
    oc = O.Context(O.Grid(nx, ny, 3e-9, 3e-9, 2e-9))
    ctx = L.Context(L.Grid2D(nx, ny, 3e-9, 3e-9, 2e-9), precision=prec)
    for mask in (0, 1, 2):
        b_ref = oc.build_h_eff(m, op, om, mask, geom=geom)
        b = ctx.build_h_eff_masked_geom(m, gp, gm, mask, geom)
        assert rel_err(b, b_ref) <= tol, mask


def test_dmi_field_flips_sign_with_d(gpu_lib):
    """tests/validation.rs:395-472 through the C ABI."""
    L = gpu_lib
    nx, ny = 9, 1
    grid = L.Grid2D(nx, ny, 1.0, 1.0, 1.0)
    m = np.zeros((nx, 3))
    for i in range(nx):
        mz = min(max((i - (nx - 1) / 2.0) * 0.01, -0.2), 0.2)
        m[i] = (np.sqrt(1 - mz * mz), 0.0, mz)
    params = L.LLGParams(alpha=0.0, dt=1e-13)
    ctx = L.Context(grid, precision=L.F64)
    bp = ctx.build_h_eff(m, params, L.Material(ms=8e5, a_ex=13e-12, dmi=1e-4))
    bm = ctx.build_h_eff(m, params, L.Material(ms=8e5, a_ex=13e-12, dmi=-1e-4))
    bxp, bxm = bp[4, 0], bm[4, 0]
    assert abs(bxp) > 0.0
    assert abs(bxp + bxm) < 1e-10 * max(abs(bxp), 1.0)


def test_torque_free_uniform_state(gpu_lib):
    """Uniform m along the easy axis with B_ext parallel: exchange and
    anisotropy give B parallel to m (closed form 2Ku/Ms + B_ext)."""
    L = gpu_lib
    nx, ny = 16, 8
    m = np.tile([0.0, 0.0, 1.0], (nx * ny, 1))
    ctx = L.Context(L.Grid2D(nx, ny, 2e-9, 2e-9, 1e-9), precision=L.F64)
    mat = L.Material(ms=5.8e5, a_ex=1.5e-11, k_u=8e5)
    b = ctx.build_h_eff(m, L.LLGParams(b_ext=(0, 0, 0.1)), mat)
    np.testing.assert_allclose(b[:, 2], 0.1 + 2 * 8e5 / 5.8e5, rtol=1e-14)
    np.testing.assert_allclose(b[:, :2], 0.0, atol=1e-14)


def test_invalid_arguments_are_rejected(gpu_lib):
    L = gpu_lib
    ctx = L.Context(L.Grid2D(4, 4, 1e-9, 1e-9, 1e-9), precision=L.F64)
    m = np.zeros((16, 3))
    with pytest.raises(L.LlgmiError) as e:
        ctx.build_h_eff_masked(m, L.LLGParams(), L.Material(ms=1.0), 7)
    assert e.value.code == L.ERR_INVALID
    with pytest.raises(L.LlgmiError) as e:
        ctx.add_demag_field_pbc(m, np.zeros_like(m), L.Material(ms=1.0, demag=True), (1 << 20) + 1, 0)
    assert e.value.code == L.ERR_UNSUPPORTED
    with pytest.raises(L.LlgmiError):
        L.Context(L.Grid2D(0, 4, 1e-9, 1e-9, 1e-9))
