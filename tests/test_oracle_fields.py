"""Pins the oracle's local field terms (CPU, no GPU) against independent
formulations rather than re-typed formulas:

* exchange (D = None, with and without a geometry mask) = minus the gradient
  of the reference's edge-based exchange energy (energy.rs:142-163:
  E = A V sum_edges |m_j - m_i|^2 / d^2), B_i = -dE/dm_i / (Ms V);
* chiral ghosts = the natural boundary condition of the interfacial DMI energy,
  2A dm/dn = -D (m_z n - (m.n) z), written in vector form;
* interior DMI field = minus the gradient of the discrete DMI energy
  D V sum_i [m_z div m - (m.grad) m_z] (central differences);
* anisotropy = minus the gradient of Ku (1 - (m.u)^2) (finite differences);
* llg_rhs: orthogonality to m and the closed form for m = x, B = z;
* reference tests: dmi_field_flips_sign_with_d (tests/validation.rs:395-472),
  energy_gradient_consistency_exchange_anisotropy (tests/validation.rs:222-318).
"""
import numpy as np
import pytest

from oracle import oracle as O


def _rand_m(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _exchange_from_energy(nx, ny, dx, dy, m, A, Ms, mask=None):
    """-dE/dm / (Ms V) of E = A V sum over material edges |dm|^2 / d^2."""
    m = m.reshape(ny, nx, 3)
    inside = np.ones((ny, nx), bool) if mask is None else mask.reshape(ny, nx).astype(bool)
    g = np.zeros_like(m)  # dE/dm / V
    # x edges (i, i+1)
    ex = inside[:, :-1] & inside[:, 1:]
    d = (m[:, 1:] - m[:, :-1]) * ex[..., None]
    g[:, 1:] += 2 * A * d / dx ** 2
    g[:, :-1] -= 2 * A * d / dx ** 2
    ey = inside[:-1, :] & inside[1:, :]
    d = (m[1:] - m[:-1]) * ey[..., None]
    g[1:] += 2 * A * d / dy ** 2
    g[:-1] -= 2 * A * d / dy ** 2
    b = -g / Ms
    b[~inside] = 0.0
    return b.reshape(-1, 3)


@pytest.mark.parametrize("shape", [(7, 5), (1, 6), (6, 1), (1, 1), (16, 9)])
@pytest.mark.parametrize("masked", [False, True])
def test_exchange_is_minus_gradient_of_edge_energy(shape, masked):
    nx, ny = shape
    dx, dy = 2e-9, 3e-9
    A, Ms = 1.3e-11, 8e5
    m = _rand_m(nx * ny, nx + 10 * ny)
    mask = None
    if masked:
        mask = np.random.default_rng(3).random(nx * ny) > 0.3
    b = np.zeros_like(m)
    O.add_exchange(O.Grid(nx, ny, dx, dy, 1e-9), m, b, O.Material(ms=Ms, a_ex=A), mask)
    # The reference's edge energy counts each pair once; its gradient gives the
    # same (2A/Ms) sum_j (m_j - m_i)/d^2 as the 5-point Laplacian with Neumann
    # (missing neighbour = centre) boundaries.
    ref = _exchange_from_energy(nx, ny, dx, dy, m, A, Ms, mask)
    scale = 2 * A / Ms / min(dx, dy) ** 2
    np.testing.assert_allclose(b, ref, rtol=0, atol=1e-12 * scale)


def _bc_ghost(m, n, eta, h):
    """Ghost value m + h dm/dn from 2A dm/dn = -D (m_z n - (m.n) z)."""
    z = np.array([0.0, 0.0, 1.0])
    dmdn = -eta * (m[2] * n - np.dot(m, n) * z)
    return m + h * dmdn


def _local_field_vector_form(nx, ny, dx, dy, m, A, Ms, D, dmi_on, mask=None):
    """Exchange with chiral ghosts + DMI, neighbours resolved in vector form."""
    m = m.reshape(ny, nx, 3)
    inside = np.ones((ny, nx), bool) if mask is None else mask.reshape(ny, nx).astype(bool)
    eta = D / (2 * A)
    out = np.zeros_like(m)
    ex, ey = np.array([1.0, 0, 0]), np.array([0, 1.0, 0])
    for j in range(ny):
        for i in range(nx):
            if not inside[j, i]:
                continue
            c = m[j, i]
            nb = {}
            for key, (di, dj, n, h) in {"xm": (-1, 0, -ex, dx), "xp": (1, 0, ex, dx),
                                        "ym": (0, -1, -ey, dy), "yp": (0, 1, ey, dy)}.items():
                ii, jj = i + di, j + dj
                axis_len = nx if di else ny
                if axis_len == 1:
                    nb[key] = ("none", c)
                elif 0 <= ii < nx and 0 <= jj < ny and inside[jj, ii]:
                    nb[key] = ("cell", m[jj, ii])
                elif 0 <= ii < nx and 0 <= jj < ny:
                    nb[key] = ("mask", _bc_ghost(c, n, eta, h))
                else:
                    nb[key] = ("edge", _bc_ghost(c, n, eta, h))
            lap = np.zeros(3)
            if nx > 1:
                lap += (nb["xp"][1] - 2 * c + nb["xm"][1]) / dx ** 2
            if ny > 1:
                lap += (nb["yp"][1] - 2 * c + nb["ym"][1]) / dy ** 2
            out[j, i] += 2 * A / Ms * lap
            if dmi_on:
                # DMI: masked neighbour is a free boundary (centre value)
                v = {k: (c if kind == "mask" else val) for k, (kind, val) in nb.items()}
                dmz_dx = (v["xp"][2] - v["xm"][2]) / (2 * dx) if nx > 1 else 0.0
                dmz_dy = (v["yp"][2] - v["ym"][2]) / (2 * dy) if ny > 1 else 0.0
                dmx_dx = (v["xp"][0] - v["xm"][0]) / (2 * dx) if nx > 1 else 0.0
                dmy_dy = (v["yp"][1] - v["ym"][1]) / (2 * dy) if ny > 1 else 0.0
                out[j, i] += 2 * D / Ms * np.array([dmz_dx, dmz_dy, -(dmx_dx + dmy_dy)])
    return out.reshape(-1, 3)


@pytest.mark.parametrize("shape", [(5, 4), (1, 5), (6, 1), (1, 1)])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("D", [3e-3, -1e-3])
def test_chiral_boundary_conditions_vector_form(shape, masked, D):
    nx, ny = shape
    dx, dy = 2e-9, 2.5e-9
    A, Ms = 1.5e-11, 5.8e5
    m = _rand_m(nx * ny, 42)
    mask = np.random.default_rng(5).random(nx * ny) > 0.35 if masked else None
    g = O.Grid(nx, ny, dx, dy, 1e-9)
    mat = O.Material(ms=Ms, a_ex=A, dmi=D)
    b = np.zeros_like(m)
    O.add_exchange(g, m, b, mat, mask)
    O.add_dmi(g, m, b, mat, mask)
    ref = _local_field_vector_form(nx, ny, dx, dy, m, A, Ms, D, True, mask)
    scale = 2 * A / Ms / min(dx, dy) ** 2
    np.testing.assert_allclose(b, ref, rtol=0, atol=1e-12 * scale)


def test_interior_dmi_is_minus_gradient_of_dmi_energy():
    nx, ny, dx, dy = 12, 10, 2e-9, 3e-9
    D, Ms, A = 2e-3, 6e5, 1e-11
    m = _rand_m(nx * ny, 8).reshape(ny, nx, 3)

    def energy(mm):
        mz = mm[..., 2]
        dmx = (mm[1:-1, 2:, 0] - mm[1:-1, :-2, 0]) / (2 * dx)
        dmy = (mm[2:, 1:-1, 1] - mm[:-2, 1:-1, 1]) / (2 * dy)
        dzx = (mm[1:-1, 2:, 2] - mm[1:-1, :-2, 2]) / (2 * dx)
        dzy = (mm[2:, 1:-1, 2] - mm[:-2, 1:-1, 2]) / (2 * dy)
        inner = mm[1:-1, 1:-1]
        return D * np.sum(mz[1:-1, 1:-1] * (dmx + dmy) - (inner[..., 0] * dzx + inner[..., 1] * dzy))

    b = np.zeros((nx * ny, 3))
    O.add_dmi(O.Grid(nx, ny, dx, dy, 1e-9), m, b, O.Material(ms=Ms, a_ex=A, dmi=D))
    b = b.reshape(ny, nx, 3)
    # dE/dm at a cell two cells away from every edge only involves interior terms
    for (j, i) in [(4, 5), (5, 6), (3, 3)]:
        for c in range(3):
            h = 1e-6
            mp, mm_ = m.copy(), m.copy()
            mp[j, i, c] += h
            mm_[j, i, c] -= h
            grad = (energy(mp) - energy(mm_)) / (2 * h)
            assert abs(-grad / Ms - b[j, i, c]) <= 1e-6 * (2 * abs(D) / Ms / dx)


def test_anisotropy_is_minus_gradient_of_energy_density():
    Ku, Ms = 4e5, 7e5
    u = np.array([0.3, -0.4, 0.866])
    u /= np.linalg.norm(u)
    m = _rand_m(20, 1)
    b = np.zeros_like(m)
    O.add_anisotropy(m, b, O.Material(ms=Ms, k_u=Ku, easy_axis=tuple(u)))
    h = 1e-7
    for k in range(20):
        for c in range(3):
            e = np.zeros(3)
            e[c] = h
            w = lambda mm: Ku * (1 - np.dot(mm, u) ** 2)
            grad = (w(m[k] + e) - w(m[k] - e)) / (2 * h)
            assert abs(-grad / Ms - b[k, c]) < 1e-7 * 2 * Ku / Ms
    # Ku = 0 or Ms = 0: no contribution
    b0 = np.zeros_like(m)
    O.add_anisotropy(m, b0, O.Material(ms=0.0, k_u=Ku))
    O.add_anisotropy(m, b0, O.Material(ms=Ms, k_u=0.0))
    assert (b0 == 0).all()


def test_zeeman_every_cell():
    b = np.zeros((6, 3))
    O.add_zeeman(b, (0.1, -0.2, 0.3))
    assert (b == np.array([0.1, -0.2, 0.3])).all()


def test_llg_rhs_closed_form_and_orthogonality():
    g, a = O.GAMMA_E, 0.3
    k = O.llg_rhs(np.array([[1.0, 0, 0]]), np.array([[0, 0, 2.0]]), g, a)[0]
    # m x B = -2y, m x (m x B) = -2z -> -g/(1+a^2) (-2y - 2a z)
    np.testing.assert_allclose(k, [0.0, 2 * g / (1 + a * a), 2 * a * g / (1 + a * a)], rtol=1e-15)
    m = _rand_m(50, 2)
    B = np.random.default_rng(4).standard_normal((50, 3))
    kk = O.llg_rhs(m, B, g, 0.1)
    assert np.abs((kk * m).sum(1)).max() < 1e-12 * np.abs(kk).max()


def test_dmi_field_flips_sign_with_d():
    nx = 9
    m = np.zeros((nx, 3))
    for i in range(nx):
        mz = min(max((i - (nx - 1) / 2) * 0.01, -0.2), 0.2)
        m[i] = (np.sqrt(1 - mz * mz), 0.0, mz)
    c = O.Context(O.Grid(nx, 1, 1.0, 1.0, 1.0))
    bp = c.build_h_eff(m, O.Params(), O.Material(ms=8e5, a_ex=13e-12, dmi=1e-4), O.FULL)
    bm = c.build_h_eff(m, O.Params(), O.Material(ms=8e5, a_ex=13e-12, dmi=-1e-4), O.FULL)
    assert abs(bp[4, 0]) > 0
    assert abs(bp[4, 0] + bm[4, 0]) < 1e-10 * max(abs(bp[4, 0]), 1.0)


def test_energy_gradient_consistency_exchange_anisotropy():
    """tests/validation.rs:222-318 with the reference's energy (energy.rs:142-175)."""
    nx, ny, d = 16, 16, 5e-9
    A, Ms, Ku = 13e-12, 8e5, 500.0
    x = (np.arange(nx) + 0.5) * d
    x0, width = 0.5 * nx * d, 5 * d
    th = 2 * np.arctan(np.exp((x - x0) / width))
    m = np.zeros((ny, nx, 3))
    m[..., 0] = np.sin(th)[None, :]
    m[..., 2] = np.cos(th)[None, :]
    m = m.reshape(-1, 3)
    V = d ** 3

    def energy(mm):
        mm = mm.reshape(ny, nx, 3)
        e = A * V * (((mm[:, 1:] - mm[:, :-1]) ** 2).sum() / d ** 2 + ((mm[1:] - mm[:-1]) ** 2).sum() / d ** 2)
        e += Ku * V * (1 - mm[..., 2] ** 2).sum()
        return e

    c = O.Context(O.Grid(nx, ny, d, d, d))
    b = c.build_h_eff(m, O.Params(), O.Material(ms=Ms, a_ex=A, k_u=Ku), O.FULL)
    idx = (ny // 2) * nx + nx // 2
    m0 = m[idx]
    a = np.array([0.37, -0.24, 0.91])
    dirn = a - np.dot(m0, a) * m0
    dirn /= np.linalg.norm(dirn)
    m1 = m0 + 1e-6 * dirn
    m1 /= np.linalg.norm(m1)
    mp = m.copy()
    mp[idx] = m1
    de_num = energy(mp) - energy(m)
    de_pred = -Ms * V * np.dot(b[idx], m1 - m0)
    scale = Ms * V * np.linalg.norm(b[idx]) * 1e-6
    assert abs(de_num - de_pred) < 5e-2 * scale
    assert abs(de_num - de_pred) < 1e-4 * scale  # much tighter than the reference's bar
