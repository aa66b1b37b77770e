"""llgmi_dev_build_h_eff_noalloc (no counterpart in the reference): the field
evaluation with its scratch in the library's per-GPU static array, against
llgmi_dev_build_h_eff on the same contexts and states."""
import numpy as np
import pytest

from conftest import DeviceBuffer

pytestmark = pytest.mark.gpu


def _film(L, nx, ny, precision, seed):
    from initial_states import smoothed_random
    es = 4 if precision == L.F32 else 8
    ctx = L.Context(L.Grid2D(nx, ny, 3e-9, 3e-9, 2e-9), precision=precision)
    m = DeviceBuffer(3 * nx * ny * es)
    ctx.dev_upload(smoothed_random(nx, ny, seed).reshape(-1, 3), m)
    return ctx, m, DeviceBuffer(3 * nx * ny * es), DeviceBuffer(3 * nx * ny * es)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_noalloc_field_equals_allocating_entry_point(gpu_lib, prec):
    """Two calls in a row, the second on a smaller grid: each equals (bitwise)
    llgmi_dev_build_h_eff on the same context and state; a third call on the
    first context again after the smaller one overwrote the array."""
    L = gpu_lib
    P = L.F32 if prec == "f32" else L.F64
    dt = np.float32 if prec == "f32" else np.float64
    mat = L.Material(ms=8e5, a_ex=1.3e-11, k_u=4e4, dmi=1.5e-3, demag=True)
    par = L.LLGParams(alpha=0.1, b_ext=(0.01, -0.02, 0.005))
    big = _film(L, 384, 256, P, 1)
    small = _film(L, 72, 40, P, 2)
    print(f"noalloc {prec}: 384x256 then 72x40", flush=True)
    for ctx, m, b_ref, b in (big, small, big):
        ctx.dev_build_h_eff(m, b_ref, par, mat, L.FieldMask.Full)     # sets up the operator
        ctx.dev_build_h_eff_noalloc(m, b, par, mat, L.FieldMask.Full)
        ctx.synchronize()
        n = 3 * ctx.grid.nx * ctx.grid.ny
        ref, got = b_ref.download(dt, n), b.download(dt, n)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, ref)
    # the local terms only (no demag, nothing of the array used)
    ctx, m, b_ref, b = small
    ctx.dev_build_h_eff(m, b_ref, par, mat, L.FieldMask.ExchAnisDmi)
    ctx.dev_build_h_eff_noalloc(m, b, par, mat, L.FieldMask.ExchAnisDmi)
    ctx.synchronize()
    n = 3 * ctx.grid.nx * ctx.grid.ny
    assert np.array_equal(b.download(dt, n), b_ref.download(dt, n))


def test_noalloc_rejects_what_it_cannot_do(gpu_lib):
    """A context whose operator is not set up, a grid whose scratch exceeds
    the array (FP64 1024 x 1024), and the multigrid demag: clear errors,
    output untouched."""
    L = gpu_lib
    mat = L.Material(ms=8e5, a_ex=1.3e-11, demag=True)
    par = L.LLGParams()
    ctx, m, _, b = _film(L, 32, 16, L.F64, 3)
    b.upload(np.full(3 * 32 * 16, 7.0))
    with pytest.raises(L.LlgmiError, match="not set up"):
        ctx.dev_build_h_eff_noalloc(m, b, par, mat, L.FieldMask.Full)
    mg = L.Material(ms=8e5, a_ex=1.3e-11, demag=True, demag_method=L.DemagMethod.PoissonMG)
    with pytest.raises(L.LlgmiError, match="FFT demag only"):
        ctx.dev_build_h_eff_noalloc(m, b, par, mg, L.FieldMask.Full)
    assert np.all(b.download(np.float64, 3 * 32 * 16) == 7.0)
    big, mb, rb, bb = _film(L, 1024, 1024, L.F64, 4)
    big.dev_build_h_eff(mb, rb, par, mat, L.FieldMask.Full)
    with pytest.raises(L.LlgmiError, match="maximum"):
        big.dev_build_h_eff_noalloc(mb, bb, par, mat, L.FieldMask.Full)
