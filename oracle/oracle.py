"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product library never
imports it.

Vector fields are numpy float64 arrays of shape (ny, nx, 3) or (n, 3), cell
index j*nx + i, as in the reference's VectorField2D (vector_field.rs:24-27).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MU0 = 4.0 * np.pi * 1e-7
GAMMA_E = 1.76085963023e11
EXCH_ANIS, EXCH_ANIS_DMI, FULL = 0, 1, 2


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double)]


class _Material(C.Structure):
    _fields_ = [("ms", C.c_double), ("a_ex", C.c_double), ("k_u", C.c_double),
                ("easy_axis", C.c_double * 3), ("has_dmi", C.c_int32), ("dmi", C.c_double),
                ("demag", C.c_int32), ("demag_method", C.c_int32)]


class _Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                ("b_ext", C.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library not built: {_LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        vp = C.c_void_p
        L.orc_add_zeeman.argtypes = [C.c_int64, dp, dp]
        L.orc_add_anisotropy.argtypes = [C.c_int64, dp, dp, C.POINTER(_Material)]
        L.orc_add_exchange.argtypes = [C.POINTER(_Grid), dp, dp, C.POINTER(_Material), u8p]
        L.orc_add_dmi.argtypes = [C.POINTER(_Grid), dp, dp, C.POINTER(_Material), u8p]
        L.orc_llg_rhs.argtypes = [C.c_int64, dp, dp, C.c_double, C.c_double, dp]
        L.orc_fft1d.argtypes = [dp, C.c_int64, C.c_int]
        L.orc_fft2.argtypes = [dp, C.c_int64, C.c_int64, C.c_int]
        L.orc_set_quad_variant.argtypes = [C.c_int]
        L.orc_kernel_components.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64,
                                            C.c_int64, C.c_double, dp]
        L.orc_build_kernel_realspace.argtypes = [C.POINTER(_Grid), C.c_int64, C.c_int64, dp, dp, dp, dp]
        L.orc_demag_new.argtypes = [C.POINTER(_Grid)]
        L.orc_demag_new.restype = vp
        L.orc_demag_new_pbc.argtypes = [C.POINTER(_Grid), C.c_int64, C.c_int64]
        L.orc_demag_new_pbc.restype = vp
        L.orc_build_kernel_realspace_pbc.argtypes = [C.POINTER(_Grid), C.c_int64, C.c_int64,
                                                     C.c_int64, C.c_int64, dp, dp, dp, dp]
        L.orc_add_demag_pbc.argtypes = [vp, dp, dp, C.POINTER(_Material), C.c_int64, C.c_int64]
        L.orc_demag_free.argtypes = [vp]
        L.orc_demag_add_field.argtypes = [vp, dp, dp, C.c_double]
        L.orc_demag_dims.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_demag_kspace.argtypes = [vp, dp, dp, dp, dp]
        L.orc_ctx_new.argtypes = [C.POINTER(_Grid)]
        L.orc_ctx_new.restype = vp
        L.orc_ctx_free.argtypes = [vp]
        L.orc_build_h_eff.argtypes = [vp, dp, dp, C.POINTER(_Params), C.POINTER(_Material),
                                      C.c_int, C.c_int, u8p]
        L.orc_add_demag.argtypes = [vp, dp, dp, C.POINTER(_Material)]
        L.orc_rk45_step.argtypes = [vp, dp, C.POINTER(_Params), C.POINTER(_Material), C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_double, dp,
                                    C.POINTER(C.c_int), dp]
        L.orc_rk45_reset.argtypes = [vp]
        L.orc_ctx_field_builds.argtypes = [vp]
        L.orc_ctx_field_builds.restype = C.c_int64
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _bind_mg(L)
        _bind_relax(L)
        _bind_hybrid(L)
        _lib = L
    return _lib


class _MGCfg(C.Structure):
    _fields_ = [("pad_xy", C.c_int64), ("n_vac_z", C.c_int64), ("v_cycles", C.c_int64),
                ("v_cycles_max", C.c_int64), ("has_tol_abs", C.c_int32), ("has_tol_rel", C.c_int32),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double), ("pre_smooth", C.c_int64),
                ("post_smooth", C.c_int64), ("smoother", C.c_int32), ("warm_start", C.c_int32),
                ("omega", C.c_double), ("sor_omega", C.c_double), ("bc", C.c_int32),
                ("stencil", C.c_int32), ("tree_theta", C.c_double), ("tree_leaf", C.c_int64),
                ("tree_max_depth", C.c_int64), ("prolong", C.c_int32), ("coarse_op", C.c_int32),
                ("iso27_alpha", C.c_double)]


MG_MAX_ST = 729


class _Stencil(C.Structure):
    _fields_ = [("center", C.c_double), ("diag", C.c_double), ("n", C.c_int32),
                ("offs", (C.c_int32 * 3) * MG_MAX_ST), ("coeffs", C.c_double * MG_MAX_ST)]


class _MGStats(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("rhs_max", C.c_double), ("tol_target", C.c_double),
                ("last_residual", C.c_double), ("diverged", C.c_int32), ("n_charges", C.c_int32)]


def _bind_mg(L):
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    i64 = C.c_int64
    sp = C.POINTER(_Stencil)
    L.orc_mg_cfg_default.argtypes = [C.POINTER(_MGCfg)]
    L.orc_mg_sanitize.argtypes = [C.POINTER(_MGCfg)]
    L.orc_stencil_seven_point.argtypes = [C.c_double] * 3 + [sp]
    L.orc_stencil_iso9_plus_z.argtypes = [C.c_double] * 3 + [sp]
    L.orc_stencil_iso27.argtypes = [C.c_double] * 4 + [sp]
    L.orc_stencil_galerkin.argtypes = [sp, C.c_int, C.c_int, C.c_int, C.c_int, sp]
    L.orc_mg_stamp_bc.argtypes = [dp, dp, i64, i64, i64]
    L.orc_mg_apply_stencil.argtypes = [sp, dp, dp, i64, i64, i64]
    L.orc_mg_smooth_wj.argtypes = [dp, dp, dp, dp, i64, i64, i64, sp, i64, C.c_double]
    L.orc_mg_smooth_rb.argtypes = [dp, dp, dp, dp, i64, i64, i64, C.c_double, C.c_double,
                                   C.c_double, i64, C.c_double]
    L.orc_mg_residual.argtypes = [dp, dp, dp, i64, i64, i64, sp]
    L.orc_mg_residual.restype = C.c_double
    L.orc_mg_restrict.argtypes = [dp, i64, i64, i64, dp, dp, i64, i64, i64]
    L.orc_mg_prolong_add.argtypes = [dp, i64, i64, i64, dp, i64, i64, i64, C.c_int]
    L.orc_mg_build_charges.argtypes = [dp, i64, i64, i64, C.c_double, C.c_double, C.c_double, dp]
    L.orc_mg_build_charges.restype = i64
    L.orc_mg_treecode_eval.argtypes = [dp, i64, dp, C.c_double, i64, C.c_double, i64, dp, i64, dp]
    L.orc_mg_treecode_bc.argtypes = [dp, dp, i64, i64, i64, C.c_double, C.c_double, C.c_double,
                                     i64, C.c_double, i64]
    L.orc_mg_dipole_bc.argtypes = [dp, dp, i64, i64, i64, C.c_double, C.c_double, C.c_double]
    L.orc_mg_new.argtypes = [C.POINTER(_Grid), C.POINTER(_MGCfg)]
    L.orc_mg_new.restype = vp
    L.orc_mg_free.argtypes = [vp]
    L.orc_mg_dims.argtypes = [vp, C.POINTER(i64)]
    L.orc_mg_levels.argtypes = [vp, C.POINTER(i64), dp]
    L.orc_mg_levels.restype = C.c_int
    L.orc_mg_level_stencil.argtypes = [vp, C.c_int, sp]
    L.orc_mg_build_rhs.argtypes = [vp, dp, C.c_double]
    L.orc_mg_solve.argtypes = [vp]
    L.orc_mg_add_b.argtypes = [vp, dp, dp]
    L.orc_mg_add_field.argtypes = [vp, dp, dp, C.c_double]
    L.orc_mg_finest.argtypes = [vp, C.c_int]
    L.orc_mg_finest.restype = dp
    L.orc_mg_stats_get.argtypes = [vp, C.POINTER(_MGStats)]
    L.orc_mg_apply_cfg.argtypes = [vp, C.POINTER(_MGCfg)]
    L.orc_ctx_set_mg_cfg.argtypes = [vp, C.POINTER(_MGCfg)]
    L.orc_ctx_mg.argtypes = [vp]
    L.orc_ctx_mg.restype = vp


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u8(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


@dataclass
class Grid:
    nx: int
    ny: int
    dx: float
    dy: float
    dz: float

    def c(self) -> _Grid:
        return _Grid(self.nx, self.ny, self.dx, self.dy, self.dz)

    @property
    def n(self) -> int:
        return self.nx * self.ny


@dataclass
class Material:
    ms: float
    a_ex: float = 0.0
    k_u: float = 0.0
    easy_axis: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    dmi: Optional[float] = None
    demag: bool = False
    demag_method: int = 0

    def c(self) -> _Material:
        return _Material(self.ms, self.a_ex, self.k_u, (C.c_double * 3)(*self.easy_axis),
                         0 if self.dmi is None else 1, 0.0 if self.dmi is None else self.dmi,
                         1 if self.demag else 0, self.demag_method)


@dataclass
class Params:
    gamma: float = GAMMA_E
    alpha: float = 0.0
    dt: float = 1e-13
    b_ext: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def c(self) -> _Params:
        return _Params(self.gamma, self.alpha, self.dt, (C.c_double * 3)(*self.b_ext))


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _mask(geom) -> Optional[np.ndarray]:
    if geom is None:
        return None
    return np.ascontiguousarray(np.asarray(geom).astype(np.uint8).reshape(-1))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


# --- individual terms -------------------------------------------------------
def add_zeeman(b: np.ndarray, b_ext) -> None:
    e = _f(b_ext)
    lib().orc_add_zeeman(b.size // 3, _dp(b), _dp(e))


def add_anisotropy(m, b: np.ndarray, mat: Material) -> None:
    m = _f(m)
    cm = mat.c()
    lib().orc_add_anisotropy(m.size // 3, _dp(m), _dp(b), C.byref(cm))


def add_exchange(grid: Grid, m, b: np.ndarray, mat: Material, mask=None) -> None:
    m = _f(m)
    mk = _mask(mask)
    g, cm = grid.c(), mat.c()
    lib().orc_add_exchange(C.byref(g), _dp(m), _dp(b), C.byref(cm), _u8(mk))


def add_dmi(grid: Grid, m, b: np.ndarray, mat: Material, mask=None) -> None:
    m = _f(m)
    mk = _mask(mask)
    g, cm = grid.c(), mat.c()
    lib().orc_add_dmi(C.byref(g), _dp(m), _dp(b), C.byref(cm), _u8(mk))


def llg_rhs(m, b, gamma: float, alpha: float) -> np.ndarray:
    m, b = _f(m), _f(b)
    k = np.zeros_like(m)
    lib().orc_llg_rhs(m.size // 3, _dp(m), _dp(b), gamma, alpha, _dp(k))
    return k


def fft1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().orc_fft1d(a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), a.size, int(inverse))
    return a


def fft2(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """x has shape (ny, nx) (row-major, x fastest)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    ny, nx = a.shape
    lib().orc_fft2(a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), nx, ny, int(inverse))
    return a


def set_quad_variant(v: int) -> None:
    """0 = the reference's quadrature rule; 1 = source faces not doubled;
    2 = sample counts truncated instead of rounded (planted mistakes that the
    tests must detect)."""
    lib().orc_set_quad_variant(int(v))


def kernel_components(dx, dy, dz, sx, sy, accuracy=10.0) -> np.ndarray:
    out = np.zeros(4)
    lib().orc_kernel_components(dx, dy, dz, sx, sy, accuracy, _dp(out))
    return out


def kernel_realspace(grid: Grid, px: int, py: int) -> np.ndarray:
    """Returns (4, py, px): kxx, kxy, kyy, kzz."""
    k = np.zeros((4, py, px))
    g = grid.c()
    lib().orc_build_kernel_realspace(C.byref(g), px, py, _dp(k[0]), _dp(k[1]), _dp(k[2]), _dp(k[3]))
    return k


def kernel_realspace_pbc(grid: Grid, px: int, py: int, pbc_x: int, pbc_y: int) -> np.ndarray:
    """Periodic-image kernels (demag_fft_uniform.rs:941-1005): (4, py, px)."""
    k = np.zeros((4, py, px))
    g = grid.c()
    lib().orc_build_kernel_realspace_pbc(C.byref(g), px, py, pbc_x, pbc_y, _dp(k[0]), _dp(k[1]),
                                         _dp(k[2]), _dp(k[3]))
    return k


class Demag:
    """Demag2D (demag_fft_uniform.rs:434-769); open boundaries unless pbc_x / pbc_y > 0."""

    def __init__(self, grid: Grid, pbc_x: int = 0, pbc_y: int = 0):
        self.grid = grid
        g = grid.c()
        self._h = lib().orc_demag_new_pbc(C.byref(g), pbc_x, pbc_y)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_demag_free(self._h)
            self._h = None

    def dims(self):
        px, py = C.c_int64(), C.c_int64()
        lib().orc_demag_dims(self._h, C.byref(px), C.byref(py))
        return px.value, py.value

    def kspace(self) -> np.ndarray:
        px, py = self.dims()
        k = np.zeros((4, py, px), dtype=np.complex128)
        p = [_dp(k[i].view(np.float64)) for i in range(4)]
        lib().orc_demag_kspace(self._h, *p)
        return k

    def add_field(self, m, b: np.ndarray, ms: float) -> None:
        m = _f(m)
        lib().orc_demag_add_field(self._h, _dp(m), _dp(b), ms)


class Context:
    """Field builder with a cached demag operator plus RK45 scratch/FSAL state."""

    def __init__(self, grid: Grid):
        self.grid = grid
        g = grid.c()
        self._h = lib().orc_ctx_new(C.byref(g))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_ctx_free(self._h)
            self._h = None

    def build_h_eff(self, m, params: Params, mat: Material, field_mask: int = FULL,
                    geom=None, use_geom: Optional[bool] = None) -> np.ndarray:
        m = _f(m)
        b = np.zeros_like(m)
        mk = _mask(geom)
        ug = (geom is not None) if use_geom is None else use_geom
        p, cm = params.c(), mat.c()
        lib().orc_build_h_eff(self._h, _dp(m), _dp(b), C.byref(p), C.byref(cm), field_mask,
                              int(ug), _u8(mk))
        return b

    def add_demag(self, m, b: np.ndarray, mat: Material) -> None:
        m = _f(m)
        cm = mat.c()
        lib().orc_add_demag(self._h, _dp(m), _dp(b), C.byref(cm))

    def add_demag_pbc(self, m, b: np.ndarray, mat: Material, pbc_x: int, pbc_y: int) -> None:
        """add_demag_field_pbc (demag.rs:70-143)."""
        m = _f(m)
        cm = mat.c()
        lib().orc_add_demag_pbc(self._h, _dp(m), _dp(b), C.byref(cm), pbc_x, pbc_y)

    def rk45_step(self, m: np.ndarray, params: Params, mat: Material, field_mask: int = FULL,
                  max_err: float = 1e-5, headroom: float = 0.8, dt_min: float = 1e-18,
                  dt_max: float = 1e-11):
        """Advances m in place; updates params.dt; returns (eps, accepted, dt_used)."""
        assert m.dtype == np.float64 and m.flags.c_contiguous
        p, cm = params.c(), mat.c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        lib().orc_rk45_step(self._h, _dp(m), C.byref(p), C.byref(cm), field_mask, max_err,
                            headroom, dt_min, dt_max, C.byref(eps), C.byref(acc), C.byref(dtu))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value

    def rk45_reset(self) -> None:
        lib().orc_rk45_reset(self._h)

    def field_builds(self) -> int:
        return int(lib().orc_ctx_field_builds(self._h))

    def set_mg_config(self, cfg) -> None:
        """MG configuration used when mat.demag_method is mg (1) or dst (2)."""
        lib().orc_ctx_set_mg_cfg(self._h, C.byref(cfg))


# --- multigrid Poisson demag (demag_poisson_mg.rs) --------------------------
SMOOTHER_WJ, SMOOTHER_RB = 0, 1
BC_ZERO, BC_DIPOLE, BC_TREECODE = 0, 1, 2
STENCIL_7, STENCIL_ISO9, STENCIL_ISO27 = 0, 1, 2
PROLONG_INJECT, PROLONG_TRILINEAR = 0, 1
COARSE_REDISCRETIZE, COARSE_GALERKIN = 0, 1


def mg_config(**kw) -> _MGCfg:
    """DemagPoissonMGConfig::default() + default operator, fields overridden by kw.
    tol_abs / tol_rel = None switch the tolerance off."""
    c = _MGCfg()
    lib().orc_mg_cfg_default(C.byref(c))
    for k, v in kw.items():
        if k in ("tol_abs", "tol_rel"):
            setattr(c, "has_" + k, 0 if v is None else 1)
            setattr(c, k, 0.0 if v is None else float(v))
        else:
            setattr(c, k, v)
    return c


def _stencil_dict(s: _Stencil) -> dict:
    n = s.n
    offs = np.array([[s.offs[q][0], s.offs[q][1], s.offs[q][2]] for q in range(n)], dtype=np.int64).reshape(n, 3)
    coeffs = np.array([s.coeffs[q] for q in range(n)])
    return {"center": s.center, "diag": s.diag, "offs": offs, "coeffs": coeffs}


def _stencil_c(d: dict) -> _Stencil:
    s = _Stencil()
    s.center, s.diag, s.n = d["center"], d["diag"], len(d["coeffs"])
    for q in range(s.n):
        for a in range(3):
            s.offs[q][a] = int(d["offs"][q][a])
        s.coeffs[q] = float(d["coeffs"][q])
    return s


def stencil(kind: int, dx: float, dy: float, dz: float, alpha: float = 0.0) -> dict:
    s = _Stencil()
    if kind == STENCIL_7:
        lib().orc_stencil_seven_point(dx, dy, dz, C.byref(s))
    elif kind == STENCIL_ISO9:
        lib().orc_stencil_iso9_plus_z(dx, dy, dz, C.byref(s))
    else:
        lib().orc_stencil_iso27(dx, dy, dz, alpha, C.byref(s))
    return _stencil_dict(s)


def galerkin(fine: dict, rx: int, ry: int, rz: int, trilinear: bool) -> dict:
    f, o = _stencil_c(fine), _Stencil()
    lib().orc_stencil_galerkin(C.byref(f), rx, ry, rz, int(trilinear), C.byref(o))
    return _stencil_dict(o)


def _box(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    assert a.ndim == 3
    return a


def mg_apply_stencil(st: dict, phi) -> np.ndarray:
    phi = _box(phi)
    nz, ny, nx = phi.shape
    out = np.zeros_like(phi)
    s = _stencil_c(st)
    lib().orc_mg_apply_stencil(C.byref(s), _dp(phi), _dp(out), nx, ny, nz)
    return out


def mg_smooth_wj(phi, rhs, bc, st: dict, iters: int, omega: float) -> np.ndarray:
    phi = _box(phi).copy()
    rhs, bc = _box(rhs), _box(bc)
    nz, ny, nx = phi.shape
    tmp = np.zeros_like(phi)
    s = _stencil_c(st)
    lib().orc_mg_smooth_wj(_dp(phi), _dp(tmp), _dp(rhs), _dp(bc), nx, ny, nz, C.byref(s), iters, omega)
    return phi


def mg_smooth_rb(phi, rhs, bc, spacing, iters: int, omega: float) -> np.ndarray:
    phi = _box(phi).copy()
    rhs, bc = _box(rhs), _box(bc)
    nz, ny, nx = phi.shape
    tmp = np.zeros_like(phi)
    dx, dy, dz = spacing
    lib().orc_mg_smooth_rb(_dp(phi), _dp(tmp), _dp(rhs), _dp(bc), nx, ny, nz, 1.0 / (dx * dx),
                           1.0 / (dy * dy), 1.0 / (dz * dz), iters, omega)
    return phi


def mg_residual(phi, rhs, st: dict):
    phi, rhs = _box(phi), _box(rhs)
    nz, ny, nx = phi.shape
    res = np.zeros_like(phi)
    s = _stencil_c(st)
    mx = lib().orc_mg_residual(_dp(phi), _dp(rhs), _dp(res), nx, ny, nz, C.byref(s))
    return res, mx


def mg_restrict(fine_res, coarse_shape):
    fr = _box(fine_res)
    fnz, fny, fnx = fr.shape
    cnz, cny, cnx = coarse_shape
    crhs = np.full(coarse_shape, np.nan)
    cphi = np.full(coarse_shape, np.nan)
    lib().orc_mg_restrict(_dp(fr), fnx, fny, fnz, _dp(crhs), _dp(cphi), cnx, cny, cnz)
    return crhs, cphi


def mg_prolong_add(coarse, fine, trilinear: bool) -> np.ndarray:
    c = _box(coarse)
    f = _box(fine).copy()
    cnz, cny, cnx = c.shape
    fnz, fny, fnx = f.shape
    lib().orc_mg_prolong_add(_dp(c), cnx, cny, cnz, _dp(f), fnx, fny, fnz, int(trilinear))
    return f


def mg_charges(rhs, spacing) -> np.ndarray:
    """(n, 4): x, y, z, q of build_charges_from_rhs."""
    rhs = _box(rhs)
    nz, ny, nx = rhs.shape
    n = lib().orc_mg_build_charges(_dp(rhs), nx, ny, nz, *spacing, None)
    out = np.zeros((n, 4))
    if n:
        lib().orc_mg_build_charges(_dp(rhs), nx, ny, nz, *spacing, _dp(out))
    return out


def treecode_eval(charges, targets, root_center, root_half, leaf=64, theta=0.6, max_depth=20):
    pq = np.ascontiguousarray(np.asarray(charges, dtype=np.float64).reshape(-1, 4))
    tg = np.ascontiguousarray(np.asarray(targets, dtype=np.float64).reshape(-1, 3))
    rc = np.ascontiguousarray(np.asarray(root_center, dtype=np.float64))
    out = np.zeros(len(tg))
    lib().orc_mg_treecode_eval(_dp(pq), len(pq), _dp(rc), root_half, leaf, theta, max_depth,
                               _dp(tg), len(tg), _dp(out))
    return out


def mg_treecode_bc(rhs, spacing, leaf=64, theta=0.6, max_depth=20) -> np.ndarray:
    rhs = _box(rhs)
    nz, ny, nx = rhs.shape
    bc = np.zeros_like(rhs)
    lib().orc_mg_treecode_bc(_dp(bc), _dp(rhs), nx, ny, nz, *spacing, leaf, theta, max_depth)
    return bc


def mg_dipole_bc(rhs, spacing) -> np.ndarray:
    rhs = _box(rhs)
    nz, ny, nx = rhs.shape
    bc = np.zeros_like(rhs)
    lib().orc_mg_dipole_bc(_dp(bc), _dp(rhs), nx, ny, nz, *spacing)
    return bc


class MG:
    """DemagPoissonMG (demag_poisson_mg.rs:1376-2098)."""

    def __init__(self, grid: Grid, cfg: Optional[_MGCfg] = None):
        self.grid = grid
        self.cfg = cfg if cfg is not None else mg_config()
        g = grid.c()
        self._h = lib().orc_mg_new(C.byref(g), C.byref(self.cfg))
        d = (C.c_int64 * 6)()
        lib().orc_mg_dims(self._h, d)
        self.px, self.py, self.pz, self.offx, self.offy, self.offz = list(d)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_mg_free(self._h)
            self._h = None

    def levels(self):
        dims = (C.c_int64 * (3 * 33))()
        sp = (C.c_double * (3 * 33))()
        n = lib().orc_mg_levels(self._h, dims, sp)
        return [((dims[3 * l + 2], dims[3 * l + 1], dims[3 * l]), (sp[3 * l], sp[3 * l + 1], sp[3 * l + 2]))
                for l in range(n)]

    def level_stencil(self, level: int) -> dict:
        s = _Stencil()
        lib().orc_mg_level_stencil(self._h, level, C.byref(s))
        return _stencil_dict(s)

    def apply_cfg(self, cfg: _MGCfg) -> None:
        self.cfg = cfg
        lib().orc_mg_apply_cfg(self._h, C.byref(cfg))

    def finest(self, which: str) -> np.ndarray:
        """Copy of the finest-level array 'phi', 'rhs', 'res' or 'bc', shape (pz, py, px)."""
        k = {"phi": 0, "rhs": 1, "res": 2, "bc": 3}[which]
        p = lib().orc_mg_finest(self._h, k)
        n = self.px * self.py * self.pz
        return np.ctypeslib.as_array(p, shape=(n,)).copy().reshape(self.pz, self.py, self.px)

    def set_finest(self, which: str, arr) -> None:
        """Overwrite a finest-level array (shape (pz, py, px))."""
        k = {"phi": 0, "rhs": 1, "res": 2, "bc": 3}[which]
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64)).reshape(-1)
        assert a.size == self.px * self.py * self.pz
        p = lib().orc_mg_finest(self._h, k)
        C.memmove(p, a.ctypes.data, a.nbytes)

    def build_rhs(self, m, ms: float) -> None:
        m = _f(m)
        lib().orc_mg_build_rhs(self._h, _dp(m), ms)

    def solve(self) -> None:
        lib().orc_mg_solve(self._h)

    def add_b(self, m, b: np.ndarray) -> None:
        m = _f(m)
        lib().orc_mg_add_b(self._h, _dp(m), _dp(b))

    def add_field(self, m, b: np.ndarray, ms: float) -> None:
        m = _f(m)
        lib().orc_mg_add_field(self._h, _dp(m), _dp(b), ms)

    def stats(self) -> dict:
        st = _MGStats()
        lib().orc_mg_stats_get(self._h, C.byref(st))
        return {"cycles": st.cycles, "rhs_max": st.rhs_max, "tol_target": st.tol_target,
                "last_residual": st.last_residual, "diverged": bool(st.diverged),
                "n_charges": st.n_charges}


# --- relaxation: RK23 relax stepper, energy, relax controller ---------------
# llg.rs:38-50, 196-436; energy.rs:36-305; relax.rs:32-808
TQ_MAX, TQ_MEAN, TQ_RMS, TQ_MEANSQ = 0, 1, 2, 3
STOP_MAX_ACCEPTED, STOP_PHASE2_DISABLED, STOP_TIGHTEN_FLOOR, STOP_GATE_NOT_SATISFIED = 0, 1, 2, 3
STAGE_NONE, STAGE_BELOW_THRESHOLD, STAGE_PLATEAU, STAGE_MAX_ACCEPTED = -1, 0, 1, 2


class RelaxSettings(C.Structure):
    """RelaxSettings (relax.rs:73-105); has_* flags stand for Option::Some."""
    _fields_ = [("max_err", C.c_double), ("headroom", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("phase1_enabled", C.c_int32),
                ("phase2_enabled", C.c_int32), ("energy_stride", C.c_int64),
                ("rel_energy_tol", C.c_double), ("torque_metric", C.c_int32),
                ("has_torque_threshold", C.c_int32), ("torque_threshold", C.c_double),
                ("torque_check_stride", C.c_int64), ("torque_plateau_rel", C.c_double),
                ("torque_plateau_abs", C.c_double), ("torque_plateau_checks", C.c_int64),
                ("torque_plateau_min_checks", C.c_int64), ("tighten_factor", C.c_double),
                ("tighten_floor", C.c_double), ("max_accepted_steps", C.c_int64),
                ("has_final_torque_gate", C.c_int32), ("has_final_torque_gate_max", C.c_int32),
                ("final_torque_gate", C.c_double), ("final_torque_gate_max", C.c_double),
                ("gate_max_extra_accepted_steps", C.c_int64), ("gate_plateau_fails", C.c_int64)]


class _RelaxReport(C.Structure):
    _fields_ = [("accepted_steps", C.c_int64), ("rejected_steps", C.c_int64),
                ("torque_checks", C.c_int64), ("torque_field_rebuilds", C.c_int64),
                ("stop_reason", C.c_int32), ("last_stage_stop", C.c_int32),
                ("has_final_torque", C.c_int32), ("has_final_torque_max", C.c_int32),
                ("final_torque", C.c_double), ("final_torque_max", C.c_double),
                ("final_max_err", C.c_double), ("final_dt", C.c_double)]


def _report_dict(r: _RelaxReport) -> dict:
    return {"accepted_steps": r.accepted_steps, "rejected_steps": r.rejected_steps,
            "torque_checks": r.torque_checks, "torque_field_rebuilds": r.torque_field_rebuilds,
            "stop_reason": r.stop_reason, "last_stage_stop": r.last_stage_stop,
            "final_torque": r.final_torque if r.has_final_torque else None,
            "final_torque_max": r.final_torque_max if r.has_final_torque_max else None,
            "final_max_err": r.final_max_err, "final_dt": r.final_dt}


def _bind_relax(L):
    dp = C.POINTER(C.c_double)
    u8p = C.POINTER(C.c_uint8)
    vp = C.c_void_p
    L.orc_llg_rhs_relax.argtypes = [C.c_int64, dp, dp, C.c_double, C.c_double, dp]
    L.orc_rk23_new.argtypes = [C.POINTER(_Grid)]
    L.orc_rk23_new.restype = vp
    L.orc_rk23_free.argtypes = [vp]
    L.orc_rk23_invalidate.argtypes = [vp]
    L.orc_rk23_last_b_eff.argtypes = [vp]
    L.orc_rk23_last_b_eff.restype = dp
    L.orc_rk23_step.argtypes = [vp, vp, dp, C.POINTER(_Params), C.POINTER(_Material), C.c_int,
                                C.c_double, C.c_double, C.c_double, C.c_double, dp,
                                C.POINTER(C.c_int), dp]
    L.orc_energy.argtypes = [vp, dp, C.POINTER(_Material), dp, u8p, dp]
    L.orc_torque_metric.argtypes = [C.c_int64, dp, dp, C.c_int]
    L.orc_torque_metric.restype = C.c_double
    L.orc_relax_settings_default.argtypes = [C.POINTER(RelaxSettings)]
    L.orc_relax_settings_mumax_like.argtypes = [C.POINTER(RelaxSettings)]
    L.orc_relax.argtypes = [vp, vp, dp, C.POINTER(_Params), C.POINTER(_Material), C.c_int,
                            C.POINTER(RelaxSettings), C.POINTER(_RelaxReport)]


def llg_rhs_relax(m, b, gamma: float, alpha: float) -> np.ndarray:
    m, b = _f(m), _f(b)
    k = np.zeros_like(m)
    lib().orc_llg_rhs_relax(m.size // 3, _dp(m), _dp(b), gamma, alpha, _dp(k))
    return k


def torque_metric(m, b, metric: int = TQ_MAX) -> float:
    m, b = _f(m), _f(b)
    return float(lib().orc_torque_metric(m.size // 3, _dp(m), _dp(b), int(metric)))


def relax_settings(**kw) -> RelaxSettings:
    """RelaxSettings::default() with fields from kw; torque_threshold /
    final_torque_gate / final_torque_gate_max = None switch that option off."""
    s = RelaxSettings()
    lib().orc_relax_settings_default(C.byref(s))
    return _apply_settings(s, kw)


def mumax_like_relax_settings(**kw) -> RelaxSettings:
    s = RelaxSettings()
    lib().orc_relax_settings_mumax_like(C.byref(s))
    return _apply_settings(s, kw)


def _apply_settings(s: RelaxSettings, kw: dict) -> RelaxSettings:
    for k, v in kw.items():
        if k in ("torque_threshold", "final_torque_gate", "final_torque_gate_max"):
            setattr(s, "has_" + k, 0 if v is None else 1)
            setattr(s, k, 0.0 if v is None else float(v))
        else:
            setattr(s, k, v)
    return s


class RK23:
    """RK23Scratch (llg.rs:196-268) for a grid, used with a Context."""

    def __init__(self, grid: Grid):
        g = grid.c()
        self.grid = grid
        self._h = lib().orc_rk23_new(C.byref(g))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rk23_free(self._h)
            self._h = None

    def invalidate(self) -> None:
        lib().orc_rk23_invalidate(self._h)

    def last_b_eff(self) -> Optional[np.ndarray]:
        p = lib().orc_rk23_last_b_eff(self._h)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(self.grid.n, 3)).copy()

    def step(self, ctx: "Context", m: np.ndarray, params: Params, mat: Material,
             field_mask: int = FULL, max_err: float = 1e-5, headroom: float = 0.8,
             dt_min: float = 1e-18, dt_max: float = 1e-11):
        """step_llg_rk23_recompute_field_masked_relax_adaptive: advances m in
        place, updates params.dt, returns (eps, accepted, dt_used)."""
        assert m.dtype == np.float64 and m.flags.c_contiguous
        p, cm = params.c(), mat.c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        lib().orc_rk23_step(ctx._h, self._h, _dp(m), C.byref(p), C.byref(cm), field_mask,
                            max_err, headroom, dt_min, dt_max, C.byref(eps), C.byref(acc),
                            C.byref(dtu))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value


ENERGY_TERMS = ("exchange", "anisotropy", "zeeman", "dmi", "demag")


def compute_energy(ctx: "Context", m, mat: Material, b_ext, geom=None) -> dict:
    """compute_energy_geom (energy.rs:80-283) -> {term: joules}."""
    m = _f(m)
    e = np.zeros(5)
    bx = _f(b_ext)
    mk = _mask(geom)
    cm = mat.c()
    lib().orc_energy(ctx._h, _dp(m), C.byref(cm), _dp(bx), _u8(mk), _dp(e))
    return dict(zip(ENERGY_TERMS, e.tolist()))


def compute_total_energy(ctx: "Context", m, mat: Material, b_ext, geom=None) -> float:
    e = compute_energy(ctx, m, mat, b_ext, geom)
    return (((e["exchange"] + e["anisotropy"]) + e["zeeman"]) + e["dmi"]) + e["demag"]


def relax_with_report(ctx: "Context", rk: RK23, m: np.ndarray, params: Params, mat: Material,
                      field_mask: int, settings: RelaxSettings) -> dict:
    """relax_with_report (relax.rs:334-774): m and params.dt in place,
    settings.max_err written back; returns the RelaxReport as a dict."""
    assert m.dtype == np.float64 and m.flags.c_contiguous
    p, cm = params.c(), mat.c()
    rep = _RelaxReport()
    lib().orc_relax(ctx._h, rk._h, _dp(m), C.byref(p), C.byref(cm), field_mask,
                    C.byref(settings), C.byref(rep))
    params.dt = p.dt
    return _report_dict(rep)


# --- hybrid MG (MG + delta-kernel correction), demag_poisson_mg.rs ---------
class HybCfg(C.Structure):
    """HybridConfig (demag_poisson_mg.rs:274-361) + BSCALE / BMARGIN."""
    _fields_ = [("enabled", C.c_int32), ("reserved", C.c_int32), ("radius_xy", C.c_int64),
                ("delta_v_cycles", C.c_int64), ("sigma_cells", C.c_double), ("bscale", C.c_double),
                ("bmargin", C.c_int64)]


def _bind_hybrid(L):
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    H = C.POINTER(HybCfg)
    L.orc_hyb_cfg_default.argtypes = [H]
    L.orc_hyb_effective_radius.argtypes = [C.POINTER(_Grid), H]
    L.orc_hyb_effective_radius.restype = C.c_int64
    L.orc_hyb_delta_kernel.argtypes = [C.POINTER(_Grid), C.POINTER(_MGCfg), H, C.c_int64, C.c_double, dp]
    L.orc_hyb_add_correction.argtypes = [C.POINTER(_Grid), C.c_int64, dp, dp, dp, C.c_double, C.c_double,
                                         C.c_int64]
    L.orc_ctx_set_hyb_cfg.argtypes = [vp, H]
    L.orc_gaussian_kernel_1d.argtypes = [C.c_double, dp, C.c_int64]
    L.orc_gaussian_kernel_1d.restype = C.c_int64
    L.orc_mg_screen_rhs.argtypes = [vp, C.c_double]
    L.orc_mg_add_b_opt.argtypes = [vp, dp, dp]


def hyb_config(**kw) -> HybCfg:
    h = HybCfg()
    lib().orc_hyb_cfg_default(C.byref(h))
    for k, v in kw.items():
        setattr(h, k, v)
    return h


def gaussian_kernel_1d(sigma: float) -> np.ndarray:
    w = np.zeros(1024)
    r = lib().orc_gaussian_kernel_1d(sigma, _dp(w), 1024)
    return w[:2 * r + 1].copy()


def hyb_effective_radius(grid: Grid, h: HybCfg) -> int:
    g = grid.c()
    return int(lib().orc_hyb_effective_radius(C.byref(g), C.byref(h)))


def hyb_delta_kernel(grid: Grid, cfg, h: HybCfg, r: int, ms: float) -> np.ndarray:
    """build_delta_kernel_impulse: (4, 2r+1, 2r+1) = xx, xy, yy, zz, [dy+r, dx+r]."""
    dk = np.zeros((4, 2 * r + 1, 2 * r + 1))
    g = grid.c()
    lib().orc_hyb_delta_kernel(C.byref(g), C.byref(cfg), C.byref(h), r, ms, _dp(dk))
    return dk


def hyb_add_correction(grid: Grid, dk: np.ndarray, m, b: np.ndarray, ms: float, bscale=0.0, bmargin=2):
    r = (dk.shape[1] - 1) // 2
    dk = np.ascontiguousarray(dk, dtype=np.float64)
    m = _f(m)
    g = grid.c()
    lib().orc_hyb_add_correction(C.byref(g), r, _dp(dk), _dp(m), _dp(b), ms, bscale, bmargin)


def _mg_screen_rhs(self, sigma: float) -> None:
    lib().orc_mg_screen_rhs(self._h, sigma)


def _mg_add_b_all(self, b: np.ndarray) -> None:
    lib().orc_mg_add_b_opt(self._h, None, _dp(b))


MG.screen_rhs = _mg_screen_rhs
MG.add_b_all = _mg_add_b_all


def _ctx_set_hyb_config(self, h: HybCfg) -> None:
    lib().orc_ctx_set_hyb_cfg(self._h, C.byref(h))


Context.set_hyb_config = _ctx_set_hyb_config
