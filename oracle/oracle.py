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
        L.orc_kernel_components.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64,
                                            C.c_int64, C.c_double, dp]
        L.orc_build_kernel_realspace.argtypes = [C.POINTER(_Grid), C.c_int64, C.c_int64, dp, dp, dp, dp]
        L.orc_demag_new.argtypes = [C.POINTER(_Grid)]
        L.orc_demag_new.restype = vp
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
        _lib = L
    return _lib


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


class Demag:
    """Demag2D (demag_fft_uniform.rs:434-769), open boundaries."""

    def __init__(self, grid: Grid):
        self.grid = grid
        g = grid.c()
        self._h = lib().orc_demag_new(C.byref(g))

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
