"""Seeded synthetic initial states and geometry masks for the benchmark workloads.

The reference has no generator for these states (its closest analogues are
initial_states.rs:27-68 and geometry_mask.rs:57-195), so they are defined
here, deterministic in the seed:

  smoothed_random   Gaussian random vectors, box-filtered over k x k cells
                    (window clamped at the edges), normalised.
  voronoi_domains   +-z domains: each cell takes the random sign of its nearest
                    of `n_points` seeded points, then per-cell angular noise of
                    relative size `noise`, normalised.
  hole_lattice_mask square lattice of circular holes (pitch, radius) as a
                    boolean material mask.

Arrays are float64 with shape (ny, nx, 3) (cell j*nx + i, AoS).
"""
from __future__ import annotations

import numpy as np


def _normalize(v: np.ndarray) -> np.ndarray:
    n = np.sqrt((v * v).sum(axis=-1, keepdims=True))
    out = np.where(n > 0, v / np.where(n > 0, n, 1.0), 0.0)
    zero = (n[..., 0] == 0)
    out[zero] = (0.0, 0.0, 1.0)
    return out


def _box_filter(a: np.ndarray, k: int) -> np.ndarray:
    """Mean over a k x k window clamped to the array (per component)."""
    ny, nx = a.shape[:2]
    r = k // 2
    c = np.cumsum(np.cumsum(np.pad(a, ((1, 0), (1, 0), (0, 0))), axis=0), axis=1)
    j0 = np.clip(np.arange(ny) - r, 0, ny)
    j1 = np.clip(np.arange(ny) + r + 1, 0, ny)
    i0 = np.clip(np.arange(nx) - r, 0, nx)
    i1 = np.clip(np.arange(nx) + r + 1, 0, nx)
    s = (c[j1][:, i1] - c[j0][:, i1] - c[j1][:, i0] + c[j0][:, i0])
    cnt = ((j1 - j0)[:, None] * (i1 - i0)[None, :])[..., None]
    return s / cnt


def smoothed_random(nx: int, ny: int, seed: int, box: int = 5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((ny, nx, 3))
    return _normalize(_box_filter(g, box))


def voronoi_domains(nx: int, ny: int, seed: int, n_points: int = 2048, noise: float = 0.1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    pts = rng.random((n_points, 2)) * np.array([nx, ny])
    sign = np.where(rng.random(n_points) < 0.5, -1.0, 1.0)
    from scipy.spatial import cKDTree
    jj, ii = np.meshgrid(np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    _, near = cKDTree(pts).query(np.stack([ii.ravel(), jj.ravel()], axis=1))
    m = np.zeros((ny * nx, 3))
    m[:, 2] = sign[near]
    m += noise * rng.standard_normal((ny * nx, 3))
    return _normalize(m).reshape(ny, nx, 3)


def hole_lattice_mask(nx: int, ny: int, dx: float, dy: float, pitch: float, radius: float) -> np.ndarray:
    """True = material.  Hole centres at ((a+1/2) pitch, (b+1/2) pitch)."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(x, y)
    fx = np.mod(X, pitch) - 0.5 * pitch
    fy = np.mod(Y, pitch) - 0.5 * pitch
    return (fx * fx + fy * fy) > radius * radius


def apply_mask_zero(m: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """m = 0 outside the material (initial_states.rs:27-34)."""
    out = m.copy()
    out[~mask.reshape(out.shape[:2])] = 0.0
    return out


def batched_films(seed: int = 64, n_films: int = 64) -> list:
    """The batched-films workload: `n_films` independent films with sizes
    drawn from {64,128,256,512} x {64,128,256}, dx = dy in U[2, 5] nm, dz = 3 nm,
    Ms in U[4e5, 1.2e6] A/m, A in U[5e-12, 2e-11] J/m, Ku in U[0, 5e5] J/m^3
    (easy axis z), D in U[0, 4e-3] J/m^2, alpha in U[0.01, 0.5], and m0 a
    Gaussian random field smoothed by three 4-neighbour averages (periodic)
    and normalised.  Each film is a dict with nx, ny, cell, material (keyword
    arguments of a material), alpha and m0 (float64, shape (ny*nx, 3))."""
    rng = np.random.default_rng(seed)
    films = []
    for _ in range(n_films):
        nx = int(rng.choice([64, 128, 256, 512]))
        ny = int(rng.choice([64, 128, 256]))
        dxy = float(rng.uniform(2e-9, 5e-9))
        mat = dict(ms=float(rng.uniform(4e5, 1.2e6)), a_ex=float(rng.uniform(5e-12, 2e-11)),
                   k_u=float(rng.uniform(0.0, 5e5)), easy_axis=(0.0, 0.0, 1.0),
                   dmi=float(rng.uniform(0.0, 4e-3)), demag=True)
        alpha = float(rng.uniform(0.01, 0.5))
        mh = rng.standard_normal((ny, nx, 3))
        for _ in range(3):
            mh = 0.25 * (np.roll(mh, 1, 0) + np.roll(mh, -1, 0) + np.roll(mh, 1, 1) + np.roll(mh, -1, 1))
        mh /= np.linalg.norm(mh, axis=2, keepdims=True)
        films.append(dict(nx=nx, ny=ny, cell=(dxy, dxy, 3e-9), material=mat, alpha=alpha,
                          m0=mh.reshape(-1, 3)))
    return films
