"""Pins of the CPU oracle's multigrid Poisson demag (demag_poisson_mg.rs,
mg_kernels.rs, mg_treecode.rs) by mathematics independent of its code:

* stencils: exactness on quadratics, zero row sums, and the closed form of the
  Galerkin operator of the 7-point stencil with injection; the trilinear
  Galerkin operator against an explicit sparse triple product R A P built with
  numpy/scipy from cell-centre geometry (np.interp);
* transfer operators against numpy reshape/kron constructions;
* smoothers, residual and V-cycle: fixed points and the converged solution
  against a scipy sparse direct solve of the same Dirichlet problem;
* RHS: closed forms for uniform films and the telescoping sum;
* gradient: exact for linear potentials;
* treecode: the theta -> 0 limit equals the direct sum; accuracy at theta 0.6;
  dipole boundary values: closed form of a point charge;
* whole path: the reference's own hybrid test records pure-MG rel-RMSE
  "~0.32" against FFT on its 32x32 random case (tests/test_demag_methods.rs:
  197-267); the oracle lands there.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O

MU0 = 4e-7 * np.pi


# --------------------------------------------------------------------------
# stencils
# --------------------------------------------------------------------------
def _quadratic_box(shape, h, a=(1.3, -0.7, 2.1)):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y, z = i * h[0], j * h[1], k * h[2]
    return a[0] * x * x + a[1] * y * y + a[2] * z * z, 2 * sum(a)


@pytest.mark.parametrize("kind,h", [
    (O.STENCIL_7, (5e-9, 4e-9, 3e-9)),
    (O.STENCIL_ISO9, (5e-9, 5e-9, 3e-9)),
    (O.STENCIL_ISO27, (5e-9, 5e-9, 3e-9)),
    (O.STENCIL_ISO27, (2e-9, 3e-9, 1e-9)),
])
def test_stencils_exact_on_quadratics(kind, h):
    st = O.stencil(kind, *h)
    assert abs(st["center"] + st["coeffs"].sum()) <= 1e-12 * abs(st["center"])
    assert st["diag"] == -st["center"]
    phi, lap = _quadratic_box((7, 7, 7), h)
    out = O.mg_apply_stencil(st, phi)
    inner = out[1:-1, 1:-1, 1:-1]
    np.testing.assert_allclose(inner, lap, rtol=1e-7)


def test_iso9_falls_back_to_seven_point_for_unequal_xy():
    a = O.stencil(O.STENCIL_ISO9, 5e-9, 4e-9, 3e-9)
    b = O.stencil(O.STENCIL_7, 5e-9, 4e-9, 3e-9)
    assert a["center"] == b["center"] and len(a["coeffs"]) == 6


def test_iso27_alpha_override_and_clamp():
    h = (5e-9, 5e-9, 3e-9)
    d = O.stencil(O.STENCIL_ISO27, *h)
    small = O.stencil(O.STENCIL_ISO27, *h, alpha=1e-30)
    huge = O.stencil(O.STENCIL_ISO27, *h, alpha=1e30)
    assert len(d["coeffs"]) == 26
    # tiny alpha: edges/corners vanish, faces -> 1/d^2 (the 7-point stencil)
    np.testing.assert_allclose(small["coeffs"][:6], O.stencil(O.STENCIL_7, *h)["coeffs"], rtol=1e-12)
    # an alpha beyond the SPD limit is clamped: face weights stay positive
    assert (huge["coeffs"][:6] > 0).all()
    # default alpha is 0.99 of the clamp
    np.testing.assert_allclose(d["coeffs"][6:], 0.99 * huge["coeffs"][6:], rtol=1e-12)


def test_galerkin_of_seven_point_with_injection_is_half_rediscretized():
    # R_a P_a = I for averaging/injection, so R A P = sum_a R_a D_a P_a and each
    # 1-D factor is [1 -2 1] / (2 h^2) in coarse units.
    h = (5e-9, 4e-9, 3e-9)
    f = O.stencil(O.STENCIL_7, *h)
    g = O.galerkin(f, 2, 2, 2, trilinear=False)
    s = {tuple(o): c for o, c in zip(g["offs"], g["coeffs"])}
    assert len(s) == 6
    for ax, hh in enumerate(h):
        for sg in (1, -1):
            o = [0, 0, 0]
            o[ax] = sg
            assert s[tuple(o)] == pytest.approx(0.5 / hh**2, rel=1e-12)
    assert g["center"] == pytest.approx(-sum(1.0 / hh**2 for hh in h), rel=1e-12)
    # ratio 1 along z: that factor keeps its fine coefficient
    g1 = O.galerkin(f, 2, 2, 1, trilinear=False)
    s1 = {tuple(o): c for o, c in zip(g1["offs"], g1["coeffs"])}
    assert s1[(0, 0, 1)] == pytest.approx(1.0 / h[2]**2, rel=1e-12)


def _interp_1d_matrix(nc, r):
    """Cell-centred piecewise-linear interpolation coarse -> fine, constant
    beyond the outermost coarse centres (np.interp)."""
    nf = nc * r
    P = np.zeros((nf, nc))
    xc = (np.arange(nc) + 0.5) * r
    for i in range(nc):
        e = np.zeros(nc)
        e[i] = 1.0
        P[:, i] = np.interp(np.arange(nf) + 0.5, xc, e)
    return P


def _avg_1d_matrix(nc, r):
    R = np.zeros((nc, nc * r))
    for i in range(nc):
        R[i, i * r:(i + 1) * r] = 1.0 / r
    return R


def _stencil_matrix(st, shape):
    """Sparse matrix of the stencil with clamped neighbour indices."""
    nz, ny, nx = shape
    n = nx * ny * nz
    idx = np.arange(n).reshape(shape)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [np.full(n, st["center"])]
    for (a, b, c), w in zip(st["offs"], st["coeffs"]):
        ii = np.clip(i + a, 0, nx - 1)
        jj = np.clip(j + b, 0, ny - 1)
        kk = np.clip(k + c, 0, nz - 1)
        rows.append(idx.ravel())
        cols.append(idx[kk, jj, ii].ravel())
        vals.append(np.full(n, w))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


@pytest.mark.parametrize("kind,h", [(O.STENCIL_ISO27, (5e-9, 5e-9, 3e-9)), (O.STENCIL_7, (4e-9, 5e-9, 2e-9))])
def test_galerkin_trilinear_matches_sparse_triple_product(kind, h):
    f = O.stencil(kind, *h)
    g = O.galerkin(f, 2, 2, 2, trilinear=True)
    nc, r = 9, 2
    P1, R1 = _interp_1d_matrix(nc, r), _avg_1d_matrix(nc, r)
    P = sp.csr_matrix(np.kron(np.kron(P1, P1), P1))   # (k, j, i) ordering
    R = sp.csr_matrix(np.kron(np.kron(R1, R1), R1))
    A = _stencil_matrix(f, (nc * r,) * 3)
    Ac = (R @ A @ P).toarray()
    c0 = (4 * nc + 4) * nc + 4
    row = Ac[c0]
    want = {}
    for kz in range(nc):
        for jy in range(nc):
            for ix in range(nc):
                v = row[(kz * nc + jy) * nc + ix]
                if abs(v) > 1e-14:
                    want[(ix - 4, jy - 4, kz - 4)] = v
    got = {tuple(o): c for o, c in zip(g["offs"], g["coeffs"])}
    got[(0, 0, 0)] = g["center"]
    assert set(got) == set(want)
    scale = abs(g["center"])
    for key in want:
        assert abs(got[key] - want[key]) <= 1e-12 * scale
    # offsets come sorted lexicographically
    offs = [tuple(o) for o in g["offs"]]
    assert offs == sorted(offs)


# --------------------------------------------------------------------------
# transfer operators
# --------------------------------------------------------------------------
def test_restrict_is_block_mean_with_zero_boundary():
    rng = np.random.default_rng(0)
    fine = rng.standard_normal((12, 10, 14))
    crhs, cphi = O.mg_restrict(fine, (6, 5, 7))
    want = fine.reshape(6, 2, 5, 2, 7, 2).mean(axis=(1, 3, 5))
    np.testing.assert_allclose(crhs[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], rtol=1e-14, atol=1e-15)
    for sl in (crhs[0], crhs[-1], crhs[:, 0], crhs[:, -1], crhs[:, :, 0], crhs[:, :, -1]):
        assert (sl == 0).all()
    assert (cphi == 0).all()
    # ratio 1 along z
    c2, _ = O.mg_restrict(fine, (12, 5, 7))
    want2 = fine.reshape(12, 5, 2, 7, 2).mean(axis=(2, 4))
    np.testing.assert_allclose(c2[1:-1, 1:-1, 1:-1], want2[1:-1, 1:-1, 1:-1], rtol=1e-14, atol=1e-15)


def test_prolong_injection_and_trilinear():
    rng = np.random.default_rng(1)
    coarse = rng.standard_normal((5, 6, 7))
    base = rng.standard_normal((10, 12, 14))
    inj = O.mg_prolong_add(coarse, base, trilinear=False)
    rep = np.repeat(np.repeat(np.repeat(coarse, 2, 0), 2, 1), 2, 2)
    # fine cells whose coarse parent touches the coarse boundary get nothing
    kc, jc, ic = np.meshgrid(np.arange(10) // 2, np.arange(12) // 2, np.arange(14) // 2, indexing="ij")
    ok = (ic > 0) & (ic + 1 < 7) & (jc > 0) & (jc + 1 < 6) & (kc > 0) & (kc + 1 < 5)
    ok[[0, -1], :, :] = False
    ok[:, [0, -1], :] = False
    ok[:, :, [0, -1]] = False
    np.testing.assert_allclose(inj, base + np.where(ok, rep, 0.0), rtol=0, atol=1e-15)

    tri = O.mg_prolong_add(coarse, base, trilinear=True)
    P = np.kron(np.kron(_interp_1d_matrix(5, 2), _interp_1d_matrix(6, 2)), _interp_1d_matrix(7, 2))
    full = (P @ coarse.ravel()).reshape(base.shape)
    interior = np.zeros_like(base, dtype=bool)
    interior[1:-1, 1:-1, 1:-1] = True
    np.testing.assert_allclose(tri, base + np.where(interior, full, 0.0), rtol=1e-14, atol=1e-14)


# --------------------------------------------------------------------------
# smoothers, residual, solver
# --------------------------------------------------------------------------
def _dirichlet_solve(st, rhs, bc):
    shape = rhs.shape
    A = _stencil_matrix(st, shape)
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    I = np.flatnonzero(inner.ravel())
    B = np.flatnonzero(~inner.ravel())
    b = rhs.ravel()[I] - A[I][:, B] @ bc.ravel()[B]
    x = bc.ravel().copy()
    x[I] = spla.spsolve(A[I][:, I].tocsc(), b)
    return x.reshape(shape)


def test_residual_matches_sparse_product():
    rng = np.random.default_rng(2)
    st = O.stencil(O.STENCIL_ISO27, 5e-9, 5e-9, 3e-9)
    phi = rng.standard_normal((6, 7, 8))
    rhs = rng.standard_normal((6, 7, 8)) * 1e17
    res, mx = O.mg_residual(phi, rhs, st)
    A = _stencil_matrix(st, phi.shape)
    want = (rhs.ravel() - A @ phi.ravel()).reshape(phi.shape)
    want[[0, -1]] = 0
    want[:, [0, -1]] = 0
    want[:, :, [0, -1]] = 0
    np.testing.assert_allclose(res, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert mx == pytest.approx(np.abs(want).max(), rel=1e-12)


def test_rb_gauss_seidel_zeroes_red_residual_and_keeps_solution():
    rng = np.random.default_rng(3)
    h = (5e-9, 4e-9, 3e-9)
    st = O.stencil(O.STENCIL_7, *h)
    shape = (6, 7, 8)
    rhs = rng.standard_normal(shape) * 1e17
    bc = np.zeros(shape)
    bc[0] = rng.standard_normal(shape[1:])
    exact = _dirichlet_solve(st, rhs, bc)
    after = O.mg_smooth_rb(exact, rhs, bc, h, 3, 1.7)
    np.testing.assert_allclose(after, exact, rtol=1e-10, atol=1e-10 * np.abs(exact).max())
    # ω = 1, one sweep: the black half-sweep leaves black residuals at zero
    phi0 = bc.copy()
    one = O.mg_smooth_rb(phi0, rhs, bc, h, 1, 1.0)
    res, _ = O.mg_residual(one, rhs, st)
    k, j, i = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    black = ((i + j + k) & 1) == 1
    black[[0, -1]] = False
    black[:, [0, -1]] = False
    black[:, :, [0, -1]] = False
    assert np.abs(res[black]).max() <= 1e-9 * np.abs(rhs).max()


def test_weighted_jacobi_sweep_matches_matrix_form():
    rng = np.random.default_rng(4)
    st = O.stencil(O.STENCIL_ISO27, 5e-9, 5e-9, 3e-9)
    shape = (6, 7, 8)
    rhs = rng.standard_normal(shape) * 1e17
    bc = rng.standard_normal(shape)
    phi = rng.standard_normal(shape)
    out = O.mg_smooth_wj(phi, rhs, bc, st, 1, 2.0 / 3.0)
    A = _stencil_matrix(st, shape)
    r = (rhs.ravel() - A @ phi.ravel()).reshape(shape)
    # (off - rhs) / diag with diag = -center and A phi = center phi + off
    want = phi + (2.0 / 3.0) * r / st["center"]
    want[[0, -1]] = bc[[0, -1]]
    want[:, [0, -1]] = bc[:, [0, -1]]
    want[:, :, [0, -1]] = bc[:, :, [0, -1]]
    np.testing.assert_allclose(out, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def _rand_m(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


@pytest.mark.parametrize("opkw", [
    dict(stencil=O.STENCIL_7, prolong=O.PROLONG_INJECT, coarse_op=O.COARSE_REDISCRETIZE, smoother=O.SMOOTHER_RB),
    dict(stencil=O.STENCIL_7, prolong=O.PROLONG_TRILINEAR, coarse_op=O.COARSE_REDISCRETIZE),
    dict(stencil=O.STENCIL_ISO27, prolong=O.PROLONG_TRILINEAR, coarse_op=O.COARSE_GALERKIN),
])
def test_solve_converges_to_sparse_direct_solution(opkw):
    g = O.Grid(10, 12, 4e-9, 4e-9, 4e-9)
    cfg = O.mg_config(pad_xy=3, n_vac_z=3, bc=O.BC_DIPOLE, tol_rel=1e-11, v_cycles=2,
                      v_cycles_max=300, **opkw)
    mg = O.MG(g, cfg)
    m = _rand_m(g.n, 5)
    mg.build_rhs(m, 8e5)
    mg.solve()
    st = mg.stats()
    assert st["last_residual"] <= st["tol_target"]
    phi = mg.finest("phi")
    exact = _dirichlet_solve(mg.level_stencil(0), mg.finest("rhs"), mg.finest("bc"))
    assert np.abs(phi - exact).max() <= 1e-8 * np.abs(exact).max()


def test_vcycle_contracts_residual():
    g = O.Grid(24, 24, 3e-9, 3e-9, 3e-9)
    cfg = O.mg_config(stencil=O.STENCIL_7, prolong=O.PROLONG_INJECT, coarse_op=O.COARSE_REDISCRETIZE,
                      smoother=O.SMOOTHER_RB, bc=O.BC_ZERO, tol_rel=None, tol_abs=None, v_cycles=1,
                      warm_start=1)
    mg = O.MG(g, cfg)
    m = _rand_m(g.n, 6)
    mg.build_rhs(m, 8e5)
    st0 = O.mg_residual(mg.finest("phi"), mg.finest("rhs"), mg.level_stencil(0))[1]
    rs = []
    for _ in range(4):
        mg.solve()
        rs.append(O.mg_residual(mg.finest("phi"), mg.finest("rhs"), mg.level_stencil(0))[1])
    assert rs[0] < 0.2 * st0
    assert all(b < 0.35 * a for a, b in zip(rs, rs[1:]))


def test_level_hierarchy_of_512_film():
    mg = O.MG(O.Grid(512, 512, 3e-9, 3e-9, 3e-9), O.mg_config(n_vac_z=16))
    assert (mg.px, mg.py, mg.pz, mg.offx, mg.offy, mg.offz) == (524, 524, 36, 6, 6, 16)
    assert [l[0] for l in mg.levels()] == [(36, 524, 524), (18, 262, 262), (9, 131, 131)]


def test_sanitize_rb_needs_seven_point_inject_rediscretize():
    c = O.mg_config(smoother=O.SMOOTHER_RB)
    O.lib().orc_mg_sanitize(c)
    assert c.smoother == O.SMOOTHER_WJ
    c = O.mg_config(smoother=O.SMOOTHER_RB, stencil=0, prolong=0, coarse_op=0)
    O.lib().orc_mg_sanitize(c)
    assert c.smoother == O.SMOOTHER_RB


# --------------------------------------------------------------------------
# RHS and gradient
# --------------------------------------------------------------------------
def test_rhs_closed_forms_for_uniform_film():
    g = O.Grid(6, 5, 2e-9, 3e-9, 1.5e-9)
    mg = O.MG(g, O.mg_config(pad_xy=2, n_vac_z=2))
    ms = 7e5
    k0, ox, oy = mg.offz, mg.offx, mg.offy
    # uniform mz: surface charges just above and below the film, nothing in it
    m = np.tile([0.0, 0.0, 1.0], (g.n, 1))
    mg.build_rhs(m, ms)
    rhs = mg.finest("rhs")
    film = (slice(oy, oy + g.ny), slice(ox, ox + g.nx))
    np.testing.assert_allclose(rhs[k0 + 1][film], -ms / g.dz, rtol=1e-15)
    np.testing.assert_allclose(rhs[k0 - 1][film], ms / g.dz, rtol=1e-15)
    assert (rhs[k0] == 0).all()
    assert abs(rhs.sum()) <= 1e-12 * np.abs(rhs).max()
    # uniform mx: line charges in the vacuum cells beside the x edges
    mg.build_rhs(np.tile([1.0, 0.0, 0.0], (g.n, 1)), ms)
    rhs = mg.finest("rhs")
    np.testing.assert_allclose(rhs[k0, oy:oy + g.ny, ox - 1], ms / g.dx, rtol=1e-15)
    np.testing.assert_allclose(rhs[k0, oy:oy + g.ny, ox + g.nx], -ms / g.dx, rtol=1e-15)
    assert np.count_nonzero(rhs) == 2 * g.ny
    # random m with holes: the discrete divergence telescopes to zero
    m = _rand_m(g.n, 7)
    m[[3, 11, 17]] = 0.0
    mg.build_rhs(m, ms)
    rhs = mg.finest("rhs")
    assert abs(rhs.sum()) <= 1e-12 * np.abs(rhs).sum()


def test_gradient_exact_for_linear_potential_and_masked():
    g = O.Grid(7, 6, 2e-9, 3e-9, 1.5e-9)
    mg = O.MG(g, O.mg_config(pad_xy=2, n_vac_z=2))
    k, j, i = np.meshgrid(np.arange(mg.pz), np.arange(mg.py), np.arange(mg.px), indexing="ij")
    a = (1.5e6, -2.0e6, 0.7e6)
    mg.set_finest("phi", a[0] * i * g.dx + a[1] * j * g.dy + a[2] * k * g.dz)
    m = _rand_m(g.n, 8)
    m[5] = 0.0
    b = np.full((g.n, 3), 0.25)
    mg.add_b(m, b)
    want = 0.25 - MU0 * np.array(a)
    mag = np.ones(g.n, dtype=bool)
    mag[5] = False
    np.testing.assert_allclose(b[mag], np.broadcast_to(want, (mag.sum(), 3)), rtol=1e-9)
    assert (b[5] == 0.25).all()


# --------------------------------------------------------------------------
# treecode / dipole boundary values
# --------------------------------------------------------------------------
def _direct_phi(ch, tg):
    d = tg[:, None, :] - ch[None, :, :3]
    r = np.sqrt((d * d).sum(-1))
    with np.errstate(divide="ignore"):
        inv = np.where(r > 0, 1.0 / r, 0.0)
    return -(inv * ch[None, :, 3]).sum(1) / (4 * np.pi)


def test_treecode_limits_direct_sum():
    rng = np.random.default_rng(9)
    ch = np.concatenate([rng.uniform(-1, 1, (3000, 3)), rng.standard_normal((3000, 1))], axis=1)
    tg = rng.uniform(-2, 2, (400, 3))
    want = _direct_phi(ch, tg)
    exact = O.treecode_eval(ch, tg, (0, 0, 0), 1.0 + 1e-12, leaf=16, theta=1e-9)
    np.testing.assert_allclose(exact, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())
    approx = O.treecode_eval(ch, tg, (0, 0, 0), 1.0 + 1e-12, leaf=16, theta=0.6)
    err = np.abs(approx - want).max() / np.abs(want).max()
    assert 0 < err < 2e-2
    # one leaf holding everything is the direct sum too
    whole = O.treecode_eval(ch, tg, (0, 0, 0), 1.0 + 1e-12, leaf=10**6, theta=0.6)
    np.testing.assert_allclose(whole, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_treecode_bc_on_box_faces_matches_direct_sum():
    rng = np.random.default_rng(10)
    shape = (8, 9, 10)
    h = (2e-9, 3e-9, 1e-9)
    rhs = np.zeros(shape)
    rhs[3:5, 2:7, 2:8] = rng.standard_normal((2, 5, 6)) * 1e17
    bc = O.mg_treecode_bc(rhs, h, leaf=8, theta=1e-9)
    ch = O.mg_charges(rhs, h)
    assert len(ch) == np.count_nonzero(rhs)
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    face = (k == 0) | (k == nz - 1) | (j == 0) | (j == ny - 1) | (i == 0) | (i == nx - 1)
    tg = np.stack([(i + 0.5 - nx / 2) * h[0], (j + 0.5 - ny / 2) * h[1], (k + 0.5 - nz / 2) * h[2]], -1)
    want = _direct_phi(ch, tg[face])
    np.testing.assert_allclose(bc[face], want, rtol=1e-10, atol=1e-12 * np.abs(want).max())
    assert (bc[~face] == 0).all()


def test_dipole_bc_of_a_charge_pair():
    shape = (9, 9, 9)
    h = (1.0, 1.0, 1.0)
    rhs = np.zeros(shape)
    rhs[4, 4, 4] = 2.0          # at the box centre: monopole only
    bc = O.mg_dipole_bc(rhs, h)
    k, j, i = np.meshgrid(*[np.arange(9)] * 3, indexing="ij")
    r = np.sqrt((i - 4.0) ** 2 + (j - 4.0) ** 2 + (k - 4.0) ** 2)
    face = (i == 0) | (i == 8) | (j == 0) | (j == 8) | (k == 0) | (k == 8)
    np.testing.assert_allclose(bc[face], -2.0 / (4 * np.pi * r[face]), rtol=1e-14)
    # +q / -q pair along x: pure dipole p = q d
    rhs[...] = 0
    rhs[4, 4, 5], rhs[4, 4, 3] = 1.0, -1.0
    bc = O.mg_dipole_bc(rhs, h)
    x = i - 4.0
    np.testing.assert_allclose(bc[face], -(2.0 * x[face]) / (4 * np.pi * r[face] ** 3), rtol=1e-12)


# --------------------------------------------------------------------------
# whole path
# --------------------------------------------------------------------------
def _xorshift_units(n, seed):
    """init_random_unit_vectors (tests/test_demag_methods.rs:146-172)."""
    M = (1 << 64) - 1
    s = seed
    out = np.zeros((n, 3))
    for c in range(n):
        v = []
        for _ in range(3):
            s ^= (s << 13) & M
            s ^= s >> 7
            s ^= (s << 17) & M
            v.append(2.0 * (s / float(M)) - 1.0)
        v = np.array(v)
        out[c] = v / max(np.sqrt((v * v).sum()), 1e-30)
    return out


def test_pure_mg_vs_fft_matches_reference_record():
    g = O.Grid(32, 32, 5e-9, 5e-9, 1e-9)
    m = _xorshift_units(g.n, 12345)
    oc = O.Context(g)
    b_fft = np.zeros_like(m)
    oc.add_demag(m, b_fft, O.Material(ms=8e5, demag=True))
    cfg = O.mg_config(warm_start=0, pad_xy=2, n_vac_z=16, v_cycles=4)
    b = np.zeros_like(m)
    O.MG(g, cfg).add_field(m, b, 8e5)
    rel = np.sqrt(((b - b_fft) ** 2).sum() / (b_fft ** 2).sum())
    assert 0.28 < rel < 0.36


def test_context_routes_mg_and_dst_and_accumulates():
    g = O.Grid(12, 10, 4e-9, 4e-9, 3e-9)
    m = _rand_m(g.n, 11)
    cfg = O.mg_config(warm_start=0, stencil=0, prolong=0, coarse_op=0, smoother=1)
    want = np.zeros_like(m)
    O.MG(g, cfg).add_field(m, want, 6e5)
    for method in (1, 2):
        oc = O.Context(g)
        oc.set_mg_config(cfg)
        b = np.full_like(m, 0.5)
        oc.add_demag(m, b, O.Material(ms=6e5, demag=True, demag_method=method))
        np.testing.assert_array_equal(b - 0.5, want + 0.5 - 0.5)


def test_thin_film_uniform_mz_gives_minus_mu0_ms():
    g = O.Grid(48, 48, 4e-9, 4e-9, 1e-9)
    m = np.tile([0.0, 0.0, 1.0], (g.n, 1))
    cfg = O.mg_config(stencil=0, prolong=0, coarse_op=0, smoother=1, tol_rel=1e-9, v_cycles_max=400)
    b = np.zeros_like(m)
    O.MG(g, cfg).add_field(m, b, 8e5)
    centre = b.reshape(48, 48, 3)[20:28, 20:28, 2]
    assert np.all(np.abs(centre / (-MU0 * 8e5) - 1) < 0.1)
