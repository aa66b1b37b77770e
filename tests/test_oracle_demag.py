"""Pins the oracle's FFT demag (demag_fft_uniform.rs) on the CPU:

* the FFT (stand-in for rustfft 6.4.1) against the DFT definition and numpy.fft;
* the quadrature self term against Aharoni's closed-form demag factors of a
  rectangular prism (J. Appl. Phys. 83, 3432 (1998)), itself checked by the
  cube value 1/3 and the sum rule Nxx + Nyy + Nzz = 1;
* far-field kernel against the point-dipole tensor;
* the parity reflection against direct quadrature at negative displacements;
* the FFT convolution against a brute-force real-space sum with the same kernel;
* reference tests: single cube Bz ~ -mu0/3 (demag_fft_uniform.rs:1261-1293),
  2x2 symmetry and factors (tests/validation.rs:478-615), demag energy >= 0
  (tests/validation.rs:617-644).
"""
import numpy as np
import pytest

from oracle import oracle as O

MU0 = O.MU0


def aharoni_n(a, b, c):
    """Demag factor along the third axis of a 2a x 2b x 2c prism."""
    r = np.sqrt(a * a + b * b + c * c)
    ab, bc, ac = np.sqrt(a * a + b * b), np.sqrt(b * b + c * c), np.sqrt(a * a + c * c)
    s = ((b * b - c * c) / (2 * b * c)) * np.log((r - a) / (r + a)) \
        + ((a * a - c * c) / (2 * a * c)) * np.log((r - b) / (r + b)) \
        + (b / (2 * c)) * np.log((ab + a) / (ab - a)) + (a / (2 * c)) * np.log((ab + b) / (ab - b)) \
        + (c / (2 * a)) * np.log((bc - b) / (bc + b)) + (c / (2 * b)) * np.log((ac - a) / (ac + a)) \
        + 2 * np.arctan(a * b / (c * r)) + (a ** 3 + b ** 3 - 2 * c ** 3) / (3 * a * b * c) \
        + ((a * a + b * b - 2 * c * c) / (3 * a * b * c)) * r + (c / (a * b)) * (ac + bc) \
        - (ab ** 3 + bc ** 3 + ac ** 3) / (3 * a * b * c)
    return s / np.pi


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 30, 64, 97, 210, 256, 1000])
def test_fft1d_matches_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    if n <= 64:
        k = np.arange(n)
        ref_dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
        np.testing.assert_allclose(ref, ref_dft, atol=1e-12 * n)
    np.testing.assert_allclose(O.fft1d(x), ref, rtol=0, atol=1e-13 * n * np.abs(x).max())
    np.testing.assert_allclose(O.fft1d(x, True), np.fft.ifft(x) * n, rtol=0, atol=1e-13 * n * np.abs(x).max())


@pytest.mark.parametrize("shape", [(4, 6), (8, 8), (10, 3), (1, 5), (64, 32)])
def test_fft2_forward_and_scaled_inverse(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    np.testing.assert_allclose(O.fft2(x), np.fft.fft2(x), atol=1e-12 * x.size)
    np.testing.assert_allclose(O.fft2(x, True), np.fft.ifft2(x), atol=1e-14)


def test_aharoni_reference_is_sound():
    assert abs(aharoni_n(1, 1, 1) - 1 / 3) < 1e-14
    for a, b, c in [(2, 2, 1.5), (1, 2, 3), (0.5, 1.5, 0.2)]:
        assert abs(aharoni_n(a, b, c) + aharoni_n(b, c, a) + aharoni_n(c, a, b) - 1) < 1e-12


@pytest.mark.parametrize("cell", [(1.0, 1.0, 1.0), (4e-9, 4e-9, 3e-9), (2e-9, 2e-9, 1e-9),
                                  (5e-9, 3e-9, 2e-9), (3e-9, 3e-9, 3e-9)])
def test_self_term_matches_aharoni(cell):
    dx, dy, dz = cell
    k = O.kernel_components(dx, dy, dz, 0, 0)
    nxx, nyy, nzz = -k[0] / MU0, -k[2] / MU0, -k[3] / MU0
    assert abs(nxx - aharoni_n(dy / 2, dz / 2, dx / 2)) < 2e-3 * nxx
    assert abs(nyy - aharoni_n(dz / 2, dx / 2, dy / 2)) < 2e-3 * nyy
    assert abs(nzz - aharoni_n(dx / 2, dy / 2, dz / 2)) < 2e-3 * nzz
    assert abs(k[1]) < 1e-12 * MU0  # no xy coupling on the self term


@pytest.mark.parametrize("s", [(40, 0), (0, 37), (25, 30), (60, 11)])
def test_far_field_is_point_dipole(s):
    dx, dy, dz = 4e-9, 3e-9, 2e-9
    sx, sy = s
    k = O.kernel_components(dx, dy, dz, sx, sy)
    r = np.array([sx * dx, sy * dy, 0.0])
    rn = np.linalg.norm(r)
    rh = r / rn
    V = dx * dy * dz
    N = V / (4 * np.pi * rn ** 3) * (3 * np.outer(rh, rh) - np.eye(3))
    tol = 5e-3 * V / (4 * np.pi * rn ** 3) * MU0
    assert abs(k[0] - MU0 * N[0, 0]) < tol
    assert abs(k[1] - MU0 * N[0, 1]) < tol
    assert abs(k[2] - MU0 * N[1, 1]) < tol
    assert abs(k[3] - MU0 * N[2, 2]) < tol


@pytest.mark.parametrize("s", [(1, 2), (3, 1), (2, 5), (0, 3), (4, 0)])
def test_reflection_parity_against_direct_quadrature(s):
    """Kernels are even (xx, yy, zz) or odd-odd (xy): the reflected base-quadrant
    values equal direct quadrature at negative displacements."""
    dx, dy, dz = 4e-9, 4e-9, 3e-9
    g = O.Grid(6, 7, dx, dy, dz)
    px, py = 12, 14
    kr = O.kernel_realspace(g, px, py)
    sx, sy = s
    for ax in (1, -1):
        for ay in (1, -1):
            direct = O.kernel_components(dx, dy, dz, ax * sx, ay * sy)
            ix, iy = (ax * sx) % px, (ay * sy) % py
            scale = np.abs(direct).max()
            for c, idx in ((0, 0), (2, 2), (3, 3)):
                assert abs(kr[c, iy, ix] - direct[idx]) <= 1e-9 * scale
            if sx != 0 and sy != 0:
                assert abs(kr[1, iy, ix] - direct[1]) <= 1e-9 * scale
            else:
                assert kr[1, iy, ix] == 0.0


def _brute_force_demag(grid, m, ms):
    """B_i = sum_j K(x_i - x_j) M_j with the oracle's real-space kernel."""
    nx, ny = grid.nx, grid.ny
    px, py = 2 * nx, 2 * ny
    kr = O.kernel_realspace(grid, px, py)
    M = ms * m.reshape(ny, nx, 3)
    B = np.zeros((ny, nx, 3))
    for j in range(ny):
        for i in range(nx):
            for jj in range(ny):
                for ii in range(nx):
                    ix, iy = (i - ii) % px, (j - jj) % py
                    kxx, kxy, kyy, kzz = kr[:, iy, ix]
                    mm = M[jj, ii]
                    B[j, i, 0] += kxx * mm[0] + kxy * mm[1]
                    B[j, i, 1] += kxy * mm[0] + kyy * mm[1]
                    B[j, i, 2] += kzz * mm[2]
    return B.reshape(-1, 3)


@pytest.mark.parametrize("shape", [(1, 1), (2, 2), (3, 5), (6, 4), (7, 1)])
def test_fft_convolution_equals_direct_sum(shape):
    nx, ny = shape
    g = O.Grid(nx, ny, 3e-9, 2e-9, 1e-9)
    m = np.random.default_rng(nx * 7 + ny).standard_normal((nx * ny, 3))
    b = np.zeros_like(m)
    O.Demag(g).add_field(m, b, 8e5)
    ref = _brute_force_demag(g, m, 8e5)
    assert np.abs(b - ref).max() <= 1e-12 * np.abs(ref).max()


def test_demag_single_cell_cube_is_close_to_minus_mu0_over_3():
    c = O.Context(O.Grid(1, 1, 1.0, 1.0, 1.0))
    b = np.zeros((1, 3))
    c.add_demag(np.array([[0.0, 0.0, 1.0]]), b, O.Material(ms=1.0, demag=True))
    assert abs(b[0, 2] + MU0 / 3) < 1e-3


def test_demag_uniform_2x2_symmetry_and_factors():
    c = O.Context(O.Grid(2, 2, 1.0, 1.0, 1.0))
    mat = O.Material(ms=1.0, demag=True)
    n = []
    for comp in range(3):
        m = np.zeros((4, 3))
        m[:, comp] = 1.0
        b = np.zeros((4, 3))
        c.add_demag(m, b, mat)
        if comp == 2:
            assert np.abs(b[:, :2]).max() < 1e-10
        assert np.ptp(b[:, comp]) < 1e-10
        assert b[0, comp] < 0
        n.append(-b[0, comp] / MU0)
    assert 0.05 < n[2] < 1.05 and 0.5 < sum(n) < 2.5
    # tighter: each cell's factors of a 2x2x1 cube array sum to ~1 (sum rule)
    assert abs(sum(n) - 1) < 2e-3


def test_demag_energy_is_nonnegative_for_uniform_state():
    g = O.Grid(32, 32, 5e-9, 5e-9, 1e-9)
    c = O.Context(g)
    m = np.tile([0.0, 0.0, 1.0], (32 * 32, 1))
    b = np.zeros_like(m)
    c.add_demag(m, b, O.Material(ms=1.0, demag=True))
    e = -0.5 * 1.0 * (m * b).sum() * g.dx * g.dy * g.dz
    assert e >= 0.0
    # thin film limit: Nzz -> 1 in the centre
    assert 0.8 < -b[16 * 32 + 16, 2] / MU0 < 1.0


# ---------------------------------------------------------------------------
# periodic images (build_kernel_realspace_pbc, demag_fft_uniform.rs:941-1005)
# ---------------------------------------------------------------------------
def _image_range(n, pbc):
    return n - 1 if pbc == 0 else n * pbc - 1


@pytest.mark.parametrize("shape,pbc", [((3, 4), (2, 0)), ((4, 3), (0, 2)), ((3, 3), (2, 3)), ((1, 2), (3, 1))])
def test_pbc_kernel_is_the_folded_image_sum(shape, pbc):
    """Each wrapped bin holds the sum of the kernel over every displacement
    |s| <= n pbc - 1 that folds onto it (computed here from the per-displacement
    quadrature with numpy), with kxy cleared on the bin axes."""
    nx, ny = shape
    g = O.Grid(nx, ny, 4e-9, 3e-9, 2e-9)
    px = nx if pbc[0] else 2 * nx
    py = ny if pbc[1] else 2 * ny
    k = O.kernel_realspace_pbc(g, px, py, *pbc)
    want = np.zeros((4, py, px))
    rx, ry = _image_range(nx, pbc[0]), _image_range(ny, pbc[1])
    for sy in range(-ry, ry + 1):
        for sx in range(-rx, rx + 1):
            want[:, sy % py, sx % px] += O.kernel_components(4e-9, 3e-9, 2e-9, sx, sy)
    want[1, :, 0] = 0.0
    want[1, 0, :] = 0.0
    assert np.abs(k - want).max() <= 1e-13 * np.abs(want).max()
    # parity of the folded bins: xx even, xy odd in each periodic direction
    for iy in range(py):
        for ix in range(px):
            jx, jy = (-ix) % px, (-iy) % py
            assert abs(k[0, iy, ix] - k[0, jy, jx]) <= 1e-12 * np.abs(k[0]).max()
            assert abs(k[1, iy, ix] + k[1, iy, jx]) <= 1e-12 * np.abs(k[0]).max()


def _brute_force_periodic(grid, m, ms, pbc):
    """B_i = sum over displacements s (|s| <= n pbc - 1 on periodic axes,
    sources at (i - s) mod n; open axes as in the open sum)."""
    nx, ny = grid.nx, grid.ny
    rx, ry = _image_range(nx, pbc[0]), _image_range(ny, pbc[1])
    M = ms * m.reshape(ny, nx, 3)
    B = np.zeros((ny, nx, 3))
    kc = {}
    for j in range(ny):
        for i in range(nx):
            for sy in range(-ry, ry + 1):
                jj = j - sy
                if pbc[1]:
                    jj %= ny
                elif not 0 <= jj < ny:
                    continue
                for sx in range(-rx, rx + 1):
                    ii = i - sx
                    if pbc[0]:
                        ii %= nx
                    elif not 0 <= ii < nx:
                        continue
                    if (sx, sy) not in kc:
                        v = O.kernel_components(grid.dx, grid.dy, grid.dz, sx, sy)
                        if sx % nx == 0 or sy % ny == 0:   # xy cleared on the bin axes
                            v = v.copy()
                            if (pbc[0] and sx % nx == 0) or sx == 0 or (pbc[1] and sy % ny == 0) or sy == 0:
                                v[1] = 0.0
                        kc[(sx, sy)] = v
                    kxx, kxy, kyy, kzz = kc[(sx, sy)]
                    mm = M[jj, ii]
                    B[j, i, 0] += kxx * mm[0] + kxy * mm[1]
                    B[j, i, 1] += kxy * mm[0] + kyy * mm[1]
                    B[j, i, 2] += kzz * mm[2]
    return B.reshape(-1, 3)


@pytest.mark.parametrize("shape,pbc", [((3, 4), (2, 0)), ((4, 3), (0, 2)), ((3, 3), (1, 2)), ((2, 1), (3, 0))])
def test_pbc_convolution_equals_direct_periodic_sum(shape, pbc):
    """The periodic operator (FFT over exactly n cells on a periodic axis)
    equals the real-space sum over every image displacement in range."""
    nx, ny = shape
    g = O.Grid(nx, ny, 3e-9, 2e-9, 1e-9)
    m = np.random.default_rng(nx * 5 + ny).standard_normal((nx * ny, 3))
    b = np.zeros_like(m)
    O.Demag(g, *pbc).add_field(m, b, 8e5)
    ref = _brute_force_periodic(g, m, 8e5, pbc)
    assert np.abs(b - ref).max() <= 1e-11 * np.abs(ref).max()


def test_pbc_translation_invariance():
    """A cyclic shift of m along a periodic axis shifts B the same way."""
    g = O.Grid(10, 6, 4e-9, 4e-9, 3e-9)
    m = np.random.default_rng(3).standard_normal((60, 3))
    d = O.Demag(g, 3, 0)
    b = np.zeros_like(m)
    d.add_field(m, b, 8e5)
    ms = np.roll(m.reshape(6, 10, 3), 4, axis=1).reshape(-1, 3)
    bs = np.zeros_like(m)
    d.add_field(ms, bs, 8e5)
    np.testing.assert_allclose(bs, np.roll(b.reshape(6, 10, 3), 4, axis=1).reshape(-1, 3),
                               rtol=0, atol=1e-12 * np.abs(b).max())


def test_pbc_thin_film_limits():
    """Periodic in x and y with many images: an out-of-plane uniform film
    approaches Bz = -mu0 Ms (Nzz -> 1), an in-plane one Bx -> 0."""
    g = O.Grid(4, 4, 5e-9, 5e-9, 2e-9)
    ms = 8e5
    d = O.Demag(g, 24, 24)
    mz = np.tile([0.0, 0.0, 1.0], (16, 1))
    b = np.zeros_like(mz)
    d.add_field(mz, b, ms)
    assert np.ptp(b[:, 2]) <= 1e-12 * abs(b[0, 2])
    assert abs(b[0, 2] / (-MU0 * ms) - 1.0) < 0.01
    mx = np.tile([1.0, 0.0, 0.0], (16, 1))
    b = np.zeros_like(mx)
    d.add_field(mx, b, ms)
    assert abs(b[0, 0]) < 0.01 * MU0 * ms
    # open boundaries for the same cell pattern are far from that limit
    bo = np.zeros_like(mz)
    O.Demag(g).add_field(mz, bo, ms)
    assert abs(bo[0, 2]) < 0.9 * MU0 * ms


def test_ctx_pbc_dispatch_falls_back_to_fft():
    """demag.rs:109-141: mg / dst with periodic axes use the FFT path."""
    g = O.Grid(6, 4, 4e-9, 4e-9, 3e-9)
    m = np.random.default_rng(8).standard_normal((24, 3))
    want = np.zeros_like(m)
    O.Demag(g, 2, 1).add_field(m, want, 8e5)
    for method in (0, 1, 2):
        ctx = O.Context(g)
        b = np.zeros_like(m)
        ctx.add_demag_pbc(m, b, O.Material(ms=8e5, demag=True, demag_method=method), 2, 1)
        np.testing.assert_array_equal(b, want)


# ---------------------------------------------------------------------------
# The quadrature's sampling rule, pinned against Newell's analytic tensor
# (A. J. Newell, W. Williams, D. J. Dunlop, J. Geophys. Res. 98, 9551 (1993)):
# the cell-averaged field of a uniformly magnetised prism from the closed-form
# f (diagonal) and g (off-diagonal) functions and their 27-point second
# differences, evaluated in 40-digit arithmetic (mpmath) where the
# differences cancel.  The accuracy-10 face-charge sum
# (mumax_like_h_from_unit_m, demag_fft_uniform.rs:1139-1242) differs from it by
# 1e-4..1.2e-3 of the largest component, a deterministic error that depends on
# the sample counts: a source face not doubled (nv *= 2, nw *= 2 dropped)
# raises it to 5e-2..9e-2 at the nearest cells, and counts truncated instead of
# rounded (mumax_n without + 0.5) change it wherever cell / max_size has a
# fractional part >= 0.5.  Each case below carries the error measured for the
# reference's rule; the test allows 25 % on it.
# ---------------------------------------------------------------------------

def _newell_f(x, y, z):
    import mpmath as mp
    x, y, z = abs(x), abs(y), abs(z)
    x2, y2, z2 = x * x, y * y, z * z
    r = mp.sqrt(x2 + y2 + z2)
    t = mp.mpf(0)
    if y != 0 and z2 != x2 and x2 + z2 != 0:
        t += (y / 2) * (z2 - x2) * mp.asinh(y / mp.sqrt(x2 + z2))
    if z != 0 and y2 != x2 and x2 + y2 != 0:
        t += (z / 2) * (y2 - x2) * mp.asinh(z / mp.sqrt(x2 + y2))
    if x != 0 and y != 0 and z != 0:
        t -= x * y * z * mp.atan(y * z / (x * r))
    return t + (2 * x2 - y2 - z2) * r / 6


def _newell_g(x, y, z):
    import mpmath as mp
    sgn = (1 if x >= 0 else -1) * (1 if y >= 0 else -1)
    x, y, z = abs(x), abs(y), abs(z)
    x2, y2, z2 = x * x, y * y, z * z
    r = mp.sqrt(x2 + y2 + z2)
    t = mp.mpf(0)
    if x != 0 and y != 0:
        if z != 0:
            t += x * y * z * mp.asinh(z / mp.sqrt(x2 + y2))
            t -= (z ** 3 / 6) * mp.atan(x * y / (z * r))
            t -= (z * y2 / 2) * mp.atan(x * z / (y * r))
            t -= (z * x2 / 2) * mp.atan(y * z / (x * r))
        t += (y / 6) * (3 * z2 - y2) * mp.asinh(x / mp.sqrt(y2 + z2))
        t += (x / 6) * (3 * z2 - x2) * mp.asinh(y / mp.sqrt(x2 + z2))
    return sgn * (t - x * y * r / 3)


def _newell_n(fun, X, Y, dx, dy, dz):
    """N = -(1 / 4 pi V) D2x D2y D2z fun at (X, Y, 0), D2 the second difference."""
    import mpmath as mp
    s = mp.mpf(0)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                c = (-2 if i == 0 else 1) * (-2 if j == 0 else 1) * (-2 if k == 0 else 1)
                s += c * fun(X + i * dx, Y + j * dy, k * dz)
    return -s / (4 * mp.pi * dx * dy * dz)


def newell_tensor(sx, sy, dx, dy, dz):
    """(Nxx, Nxy, Nyy, Nzz) between two dx x dy x dz cells sx, sy cells apart."""
    import mpmath as mp
    with mp.workdps(40):
        dx, dy, dz = mp.mpf(dx), mp.mpf(dy), mp.mpf(dz)
        X, Y = sx * dx, sy * dy
        nxx = _newell_n(_newell_f, X, Y, dx, dy, dz)
        nyy = _newell_n(lambda a, b, c: _newell_f(b, a, c), X, Y, dx, dy, dz)
        nzz = _newell_n(lambda a, b, c: _newell_f(c, b, a), X, Y, dx, dy, dz)
        nxy = _newell_n(_newell_g, X, Y, dx, dy, dz)
        return np.array([float(nxx), float(nxy), float(nyy), float(nzz)])


def test_newell_tensor_is_sound():
    """Cube self term 1/3, trace 1 on the self term for any prism, Aharoni's
    prism factors, the point-dipole far field, Nxy = 0 on the axes."""
    assert np.allclose(newell_tensor(0, 0, 1, 1, 1), [1 / 3, 0, 1 / 3, 1 / 3], atol=1e-15)
    for cell in [(4, 4, 3), (5, 3, 2), (2, 2, 1)]:
        n = newell_tensor(0, 0, *cell)
        assert abs(n[0] + n[2] + n[3] - 1) < 1e-14
        dx, dy, dz = cell
        assert abs(n[0] - aharoni_n(dy / 2, dz / 2, dx / 2)) < 1e-13
        assert abs(n[3] - aharoni_n(dx / 2, dy / 2, dz / 2)) < 1e-13
        assert n[1] == 0.0 and newell_tensor(3, 0, *cell)[1] == 0.0
    dx, dy, dz = 4.0, 3.0, 2.0
    sx, sy = 60, 45
    r = np.array([sx * dx, sy * dy, 0.0])
    rn = np.linalg.norm(r)
    rh = r / rn
    nd = -(dx * dy * dz) / (4 * np.pi * rn ** 3) * (3 * np.outer(rh, rh) - np.eye(3))
    n = newell_tensor(sx, sy, dx, dy, dz)
    assert np.allclose(n, [nd[0, 0], nd[0, 1], nd[1, 1], nd[2, 2]], rtol=2e-3, atol=0)


# (cell in nm, displacement) -> max |N_quad - N_newell| / max |N_newell| measured for the reference's rule
QUAD_VS_NEWELL = {
    ((4, 4, 3), (0, 0)): 7.91e-4, ((4, 4, 3), (1, 0)): 8.42e-4, ((4, 4, 3), (1, 1)): 8.56e-4,
    ((4, 4, 3), (2, 0)): 2.15e-4, ((4, 4, 3), (2, 1)): 2.20e-4, ((4, 4, 3), (1, 2)): 2.20e-4,
    ((4, 4, 3), (3, 2)): 3.42e-4,
    ((2, 2, 1), (0, 0)): 8.39e-4, ((2, 2, 1), (1, 0)): 6.59e-4, ((2, 2, 1), (2, 1)): 1.69e-4,
    ((5, 3, 2), (0, 0)): 8.30e-4, ((5, 3, 2), (0, 1)): 8.09e-4, ((5, 3, 2), (1, 2)): 1.82e-4,
    ((5, 3, 2), (3, 2)): 2.83e-4,
    ((3, 3, 3), (0, 0)): 5.86e-4, ((3, 3, 3), (1, 1)): 1.16e-3,
    ((4, 3, 2), (0, 0)): 8.18e-4, ((4, 3, 2), (2, 0)): 2.75e-4, ((4, 3, 2), (1, 2)): 2.16e-4,
    ((4, 3, 2), (3, 2)): 9.27e-4,
}


def _quad_err(cell_nm, disp):
    cell = tuple(c * 1e-9 for c in cell_nm)
    nq = -np.asarray(O.kernel_components(*cell, *disp)) / MU0
    nw = newell_tensor(*disp, *cell_nm)
    return float(np.abs(nq - nw).max() / np.abs(nw).max())


@pytest.mark.parametrize("case", sorted(QUAD_VS_NEWELL))
def test_quadrature_error_against_newell(case):
    """The oracle's accuracy-10 kernel differs from the analytic tensor by the
    recorded amount (within 25 %), and by at most 1.2e-3 anywhere."""
    err = _quad_err(*case)
    ref = QUAD_VS_NEWELL[case]
    assert 0.8 * ref <= err <= 1.25 * ref, (case, err)
    assert err <= 1.2e-3


def test_planted_quadrature_mistakes_fail_the_newell_bands():
    """With each planted mistake the oracle leaves the bands above: an
    undoubled source face on every near-field case, truncated sample counts
    on every case where a count has a fractional part >= 0.5."""
    try:
        O.set_quad_variant(1)
        for case, ref in QUAD_VS_NEWELL.items():
            assert _quad_err(*case) > 1.25 * ref, case
        O.set_quad_variant(2)
        moved = [c for c, ref in QUAD_VS_NEWELL.items() if not 0.8 * ref <= _quad_err(*c) <= 1.25 * ref]
        assert set(moved) == {((4, 4, 3), (2, 0)), ((4, 4, 3), (1, 2)), ((5, 3, 2), (1, 2)),
                              ((5, 3, 2), (3, 2)), ((4, 3, 2), (2, 0)), ((4, 3, 2), (1, 2)),
                              ((4, 3, 2), (3, 2))}, moved
    finally:
        O.set_quad_variant(0)
