"""Python binding of libllgmi.so (the C ABI in include/llgmi.h) via ctypes.

Mirrors the Rust API of llg-sim's hot path by name:
  Grid2D (grid.rs:5-31), Material / LLGParams (params.rs:94-144),
  FieldMask (effective_field/mod.rs:37-45),
  Context.build_h_eff_masked[_geom] (mod.rs:48-101),
  Context.add_demag_field_pbc (demag.rs:70-143),
  Context.step_llg_rk45_recompute_field_masked_adaptive (llg.rs:795-1026).

Host-side methods take numpy float64 arrays of shape (n, 3) or (ny, nx, 3)
(AoS, cell j*nx + i).  Device methods take torch tensors (or raw pointers)
in SoA layout (3, ny, nx) in the context precision.

The library is required: importing works without it, but creating a context
raises if libllgmi.so is missing, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libllgmi.so")

MU0 = 4.0 * np.pi * 1e-7
GAMMA_E_RAD_PER_S_T = 1.76085963023e11

OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED, ERR_ALLOC = 0, 1, 2, 3, 4
F32, F64 = 0, 1

(K_FFT_ROW_FWD, K_FFT_COL, K_FFT_ROW_INV, K_LOCAL_FIELD, K_RK45_STAGE, K_RK45_ERR, K_MG, K_MG_BC,
 K_RELAX) = range(9)
K_COUNT = 9
KERNEL_CLASS_NAMES = {
    K_FFT_ROW_FWD: "fft_row_fwd",
    K_FFT_COL: "fft_col_mul",
    K_FFT_ROW_INV: "fft_row_inv",
    K_LOCAL_FIELD: "local_field",
    K_RK45_STAGE: "rk45_stage",
    K_RK45_ERR: "rk45_err",
    K_MG: "mg",
    K_MG_BC: "mg_bc",
    K_RELAX: "relax",
}


class FieldMask:
    """FieldMask (effective_field/mod.rs:37-45)."""
    ExchAnis = 0
    ExchAnisDmi = 1
    Full = 2


class DemagMethod:
    """DemagMethod (params.rs:67-92)."""
    FftUniform = 0
    PoissonMG = 1
    PoissonDst = 2


class LlgmiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"llgmi error {code}: {msg}")
        self.code = code


class _Grid(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("dx", C.c_double), ("dy", C.c_double),
                ("dz", C.c_double)]


class _Material(C.Structure):
    _fields_ = [("ms", C.c_double), ("a_ex", C.c_double), ("k_u", C.c_double),
                ("easy_axis", C.c_double * 3), ("has_dmi", C.c_int32), ("demag", C.c_int32),
                ("dmi", C.c_double), ("demag_method", C.c_int32), ("reserved", C.c_int32)]


class _Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("alpha", C.c_double), ("dt", C.c_double),
                ("b_ext", C.c_double * 3)]


@dataclass
class Grid2D:
    nx: int
    ny: int
    dx: float
    dy: float
    dz: float

    def n_cells(self) -> int:
        return self.nx * self.ny

    def cell_volume(self) -> float:
        return self.dx * self.dy * self.dz

    def idx(self, i: int, j: int) -> int:
        return j * self.nx + i

    def _c(self) -> _Grid:
        return _Grid(self.nx, self.ny, self.dx, self.dy, self.dz)


@dataclass
class Material:
    ms: float
    a_ex: float = 0.0
    k_u: float = 0.0
    easy_axis: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    dmi: Optional[float] = None
    demag: bool = False
    demag_method: int = DemagMethod.FftUniform

    def _c(self) -> _Material:
        return _Material(self.ms, self.a_ex, self.k_u, (C.c_double * 3)(*self.easy_axis),
                         0 if self.dmi is None else 1, 1 if self.demag else 0,
                         0.0 if self.dmi is None else float(self.dmi), int(self.demag_method), 0)


@dataclass
class LLGParams:
    gamma: float = GAMMA_E_RAD_PER_S_T
    alpha: float = 0.0
    dt: float = 1e-13
    b_ext: Tuple[float, float, float] = (0.0, 0.0, 0.0)

    def _c(self) -> _Params:
        return _Params(self.gamma, self.alpha, self.dt, (C.c_double * 3)(*self.b_ext))


_lib = None

# Every symbol include/llgmi.h declares (checked by the CPU test suite).
EXPORTED = [
    "llgmi_last_error", "llgmi_version", "llgmi_device_count", "llgmi_ctx_create",
    "llgmi_ctx_destroy", "llgmi_ctx_stream", "llgmi_ctx_synchronize",
    "llgmi_build_h_eff_masked", "llgmi_build_h_eff_masked_geom", "llgmi_build_h_eff",
    "llgmi_add_demag_field_pbc", "llgmi_compute_demag_field_pbc", "llgmi_add_demag_field",
    "llgmi_rk45_step", "llgmi_rk45_step_adaptive", "llgmi_rk45_reset",
    "llgmi_dev_build_h_eff", "llgmi_dev_build_h_eff_noalloc", "llgmi_dev_add_demag_field_pbc", "llgmi_dev_rk45_step",
    "llgmi_dev_rk45_step_batch", "llgmi_dev_rk45_run_batch", "llgmi_dev_rk45_run",
    "llgmi_dev_upload", "llgmi_dev_download", "llgmi_profile_enable", "llgmi_profile_reset",
    "llgmi_profile_read", "llgmi_profile_fft_passes", "llgmi_mg_config_default", "llgmi_mg_config_from_env",
    "llgmi_ctx_set_mg_config", "llgmi_add_demag_field_poisson_mg",
    "llgmi_compute_demag_field_poisson_mg", "llgmi_mg_last_stats",
    "llgmi_relax_settings_default", "llgmi_relax_settings_mumax_like", "llgmi_rk23_relax_step",
    "llgmi_rk23_invalidate", "llgmi_rk23_last_b_eff", "llgmi_compute_energy_geom",
    "llgmi_relax_with_report", "llgmi_dev_rk23_relax_step", "llgmi_dev_compute_energy_geom",
    "llgmi_dev_relax_with_report", "llgmi_write_ovf2_text", "llgmi_write_ovf2_binary4",
    "llgmi_dev_write_ovf2", "llgmi_ovf_total_sim_time_line", "llgmi_demag_cache_path",
    "llgmi_demag_cache_save", "llgmi_demag_cache_load", "llgmi_mg_hybrid_config_default",
    "llgmi_mg_hybrid_config_from_env", "llgmi_ctx_set_mg_hybrid_config",
    "llgmi_mg_box", "llgmi_mg_phi_read", "llgmi_mg_phi_write", "llgmi_dev_mask_cells",
    "llgmi_ws_create", "llgmi_ws_destroy", "llgmi_ws_param_count", "llgmi_ws_get_params",
    "llgmi_ws_set_params", "llgmi_ws_get_grads", "llgmi_ws_forward", "llgmi_ws_loss_grad",
    "llgmi_ws_adam_step", "llgmi_ws_train_step", "llgmi_ws_pack_input", "llgmi_ws_pack_phi",
    "llgmi_ws_unpack_phi",
]


class WSShape(C.Structure):
    """llgmi_ws_shape (56 bytes)."""
    _fields_ = [("batch", C.c_int64), ("height", C.c_int64), ("width", C.c_int64),
                ("in_channels", C.c_int32), ("out_channels", C.c_int32),
                ("base_channels", C.c_int32), ("reserved", C.c_int32),
                ("valid_width", C.c_int64), ("valid_height", C.c_int64)]


class MGConfig(C.Structure):
    """llgmi_mg_config: DemagPoissonMGConfig (demag_poisson_mg.rs:115-185) with the
    operator choice (MGOperatorSettings, :727-804)."""
    _fields_ = [("pad_xy", C.c_int64), ("n_vac_z", C.c_int64), ("v_cycles", C.c_int64),
                ("v_cycles_max", C.c_int64), ("has_tol_abs", C.c_int32), ("has_tol_rel", C.c_int32),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double), ("pre_smooth", C.c_int64),
                ("post_smooth", C.c_int64), ("smoother", C.c_int32), ("warm_start", C.c_int32),
                ("omega", C.c_double), ("sor_omega", C.c_double), ("bc", C.c_int32),
                ("stencil", C.c_int32), ("tree_theta", C.c_double), ("tree_leaf", C.c_int64),
                ("tree_max_depth", C.c_int64), ("prolong", C.c_int32), ("coarse_op", C.c_int32),
                ("iso27_alpha", C.c_double)]


class MGStats(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("rhs_max", C.c_double), ("tol_target", C.c_double),
                ("last_residual", C.c_double), ("diverged", C.c_int32), ("n_charges", C.c_int32),
                ("tree_rebuilt", C.c_int32), ("levels", C.c_int32), ("box", C.c_int64 * 3)]

    def as_dict(self) -> dict:
        return {"cycles": self.cycles, "rhs_max": self.rhs_max, "tol_target": self.tol_target,
                "last_residual": self.last_residual, "diverged": bool(self.diverged),
                "n_charges": self.n_charges, "tree_rebuilt": bool(self.tree_rebuilt),
                "levels": self.levels, "box": tuple(self.box)}


SMOOTHER_WJ, SMOOTHER_RB = 0, 1
BC_ZERO, BC_DIPOLE, BC_TREECODE = 0, 1, 2
STENCIL_7, STENCIL_ISO9, STENCIL_ISO27 = 0, 1, 2
PROLONG_INJECT, PROLONG_TRILINEAR = 0, 1
COARSE_REDISCRETIZE, COARSE_GALERKIN = 0, 1


class MGHybridConfig(C.Structure):
    """llgmi_mg_hybrid_config: HybridConfig (demag_poisson_mg.rs:274-361) + BSCALE / BMARGIN."""
    _fields_ = [("enabled", C.c_int32), ("reserved", C.c_int32), ("radius_xy", C.c_int64),
                ("delta_v_cycles", C.c_int64), ("sigma_cells", C.c_double), ("bscale", C.c_double),
                ("bmargin", C.c_int64)]


def mg_hybrid_config(**kw) -> MGHybridConfig:
    """HybridConfig::default() (disabled) with fields from kw."""
    c = MGHybridConfig()
    _check(lib().llgmi_mg_hybrid_config_default(C.byref(c)))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def mg_hybrid_config_from_env() -> MGHybridConfig:
    c = MGHybridConfig()
    _check(lib().llgmi_mg_hybrid_config_from_env(C.byref(c)))
    return c


def mg_config(**kw) -> MGConfig:
    """DemagPoissonMGConfig::default() with the default operator, fields from kw.
    tol_abs / tol_rel = None switch that tolerance off."""
    c = MGConfig()
    _check(lib().llgmi_mg_config_default(C.byref(c)))
    for k, v in kw.items():
        if k in ("tol_abs", "tol_rel"):
            setattr(c, "has_" + k, 0 if v is None else 1)
            setattr(c, k, 0.0 if v is None else float(v))
        else:
            setattr(c, k, v)
    return c


def mg_config_from_env() -> MGConfig:
    """DemagPoissonMGConfig::from_env (:188-267) + MGOperatorSettings::from_env."""
    c = MGConfig()
    _check(lib().llgmi_mg_config_from_env(C.byref(c)))
    return c


# ---- relaxation (relax.rs, energy.rs, llg.rs:196-436) ----
TORQUE_MAX, TORQUE_MEAN, TORQUE_RMS, TORQUE_MEANSQ = 0, 1, 2, 3
(STOP_MAX_ACCEPTED_STEPS, STOP_PHASE2_DISABLED, STOP_TIGHTEN_FLOOR_REACHED,
 STOP_TORQUE_GATE_NOT_SATISFIED) = 0, 1, 2, 3
STAGE_NONE, STAGE_BELOW_THRESHOLD, STAGE_PLATEAU, STAGE_MAX_ACCEPTED_STEPS = -1, 0, 1, 2


class RelaxSettings(C.Structure):
    """llgmi_relax_settings: RelaxSettings (relax.rs:73-105); has_* = 0 is None."""
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


class RelaxReport(C.Structure):
    """llgmi_relax_report: RelaxReport (relax.rs:55-71)."""
    _fields_ = [("accepted_steps", C.c_int64), ("rejected_steps", C.c_int64),
                ("torque_checks", C.c_int64), ("torque_field_rebuilds", C.c_int64),
                ("stop_reason", C.c_int32), ("last_stage_stop", C.c_int32),
                ("has_final_torque", C.c_int32), ("has_final_torque_max", C.c_int32),
                ("final_torque", C.c_double), ("final_torque_max", C.c_double),
                ("final_max_err", C.c_double), ("final_dt", C.c_double)]

    def as_dict(self) -> dict:
        return {"accepted_steps": self.accepted_steps, "rejected_steps": self.rejected_steps,
                "torque_checks": self.torque_checks,
                "torque_field_rebuilds": self.torque_field_rebuilds,
                "stop_reason": self.stop_reason, "last_stage_stop": self.last_stage_stop,
                "final_torque": self.final_torque if self.has_final_torque else None,
                "final_torque_max": self.final_torque_max if self.has_final_torque_max else None,
                "final_max_err": self.final_max_err, "final_dt": self.final_dt}


class Energy(C.Structure):
    """llgmi_energy: EnergyBreakdown (energy.rs:21-34), joules."""
    _fields_ = [("exchange", C.c_double), ("anisotropy", C.c_double), ("zeeman", C.c_double),
                ("dmi", C.c_double), ("demag", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}

    def total(self) -> float:
        return self.exchange + self.anisotropy + self.zeeman + self.dmi + self.demag


def _settings(fill, kw) -> RelaxSettings:
    s = RelaxSettings()
    _check(fill(C.byref(s)))
    for k, v in kw.items():
        if k in ("torque_threshold", "final_torque_gate", "final_torque_gate_max"):
            setattr(s, "has_" + k, 0 if v is None else 1)
            setattr(s, k, 0.0 if v is None else float(v))
        else:
            setattr(s, k, v)
    return s


def relax_settings(**kw) -> RelaxSettings:
    """RelaxSettings::default() with fields from kw; None switches an Option field off."""
    return _settings(lib().llgmi_relax_settings_default, kw)


def mumax_like_relax_settings(**kw) -> RelaxSettings:
    """mumax_like_relax_settings (relax.rs:789-808) with fields from kw."""
    return _settings(lib().llgmi_relax_settings_mumax_like, kw)


class OvfMeta(C.Structure):
    """llgmi_ovf_meta: OvfMeta (ovf.rs:17-46); None fields take magnetization()."""
    _fields_ = [("title", C.c_char_p), ("desc_lines", C.POINTER(C.c_char_p)), ("n_desc", C.c_int32),
                ("reserved", C.c_int32), ("valuelabels", C.c_char_p * 3), ("valueunits", C.c_char_p * 3)]


def ovf_meta(title=None, desc_lines=(), valuelabels=None, valueunits=None) -> OvfMeta:
    m = OvfMeta()
    m.title = None if title is None else title.encode()
    lines = [d.encode() for d in desc_lines]
    arr = (C.c_char_p * max(len(lines), 1))(*lines)
    m._keep = arr                                    # keeps the strings alive with the struct
    m.desc_lines = C.cast(arr, C.POINTER(C.c_char_p))
    m.n_desc = len(lines)
    for q in range(3):
        m.valuelabels[q] = None if valuelabels is None else valuelabels[q].encode()
        m.valueunits[q] = None if valueunits is None else valueunits[q].encode()
    return m


def ovf_total_sim_time_line(t_s: float) -> str:
    """OvfMeta::with_total_sim_time's desc line (ovf.rs:36-41)."""
    buf = C.create_string_buffer(128)
    _check(lib().llgmi_ovf_total_sim_time_line(t_s, buf, 128))
    return buf.value.decode()


def _aos(grid: "Grid2D", m) -> np.ndarray:
    a = _f64(m).reshape(-1, 3)
    assert a.shape[0] == grid.n_cells()
    return a


def write_ovf2_rectangular_text(path: str, grid: "Grid2D", m, meta: Optional[OvfMeta] = None) -> None:
    """ovf.rs:54-159."""
    a = _aos(grid, m)
    g = grid._c()
    _check(lib().llgmi_write_ovf2_text(path.encode(), C.byref(g), _dp(a),
                                       None if meta is None else C.byref(meta)))


def write_ovf2_rectangular_binary4(path: str, grid: "Grid2D", m, meta: Optional[OvfMeta] = None) -> None:
    """ovf.rs:161-262."""
    a = _aos(grid, m)
    g = grid._c()
    _check(lib().llgmi_write_ovf2_binary4(path.encode(), C.byref(g), _dp(a),
                                          None if meta is None else C.byref(meta)))


def demag_cache_path(grid: "Grid2D", pbc_x: int = 0, pbc_y: int = 0) -> str:
    """demag_cache_path (demag_fft_uniform.rs:262-280)."""
    buf = C.create_string_buffer(512)
    g = grid._c()
    _check(lib().llgmi_demag_cache_path(C.byref(g), pbc_x, pbc_y, buf, 512))
    return buf.value.decode()


def lib():
    """Loads libllgmi.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libllgmi.so not found at {LIB_PATH}: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    dp, ip, vp, u8p = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p, C.POINTER(C.c_uint8)
    G, M, P = C.POINTER(_Grid), C.POINTER(_Material), C.POINTER(_Params)
    L.llgmi_last_error.restype = C.c_char_p
    L.llgmi_device_count.argtypes = [ip]
    L.llgmi_ctx_create.argtypes = [G, C.c_int, C.c_int, C.POINTER(vp)]
    L.llgmi_ctx_destroy.argtypes = [vp]
    L.llgmi_ctx_stream.argtypes = [vp, C.POINTER(vp)]
    L.llgmi_ctx_synchronize.argtypes = [vp]
    L.llgmi_build_h_eff_masked.argtypes = [vp, dp, dp, P, M, C.c_int]
    L.llgmi_build_h_eff_masked_geom.argtypes = [vp, dp, dp, P, M, C.c_int, u8p]
    L.llgmi_build_h_eff.argtypes = [vp, dp, dp, P, M]
    L.llgmi_add_demag_field_pbc.argtypes = [vp, dp, dp, M, C.c_uint64, C.c_uint64]
    L.llgmi_compute_demag_field_pbc.argtypes = [vp, dp, dp, M, C.c_uint64, C.c_uint64]
    L.llgmi_add_demag_field.argtypes = [vp, dp, dp, M]
    step_args = [vp, dp, P, M, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, dp, ip, dp]
    L.llgmi_rk45_step.argtypes = step_args
    L.llgmi_rk45_step_adaptive.argtypes = [vp, dp, P, M, C.c_double, C.c_double, C.c_double,
                                           C.c_double, dp, ip, dp]
    L.llgmi_rk45_reset.argtypes = [vp]
    L.llgmi_dev_build_h_eff.argtypes = [vp, vp, vp, P, M, C.c_int, vp]
    L.llgmi_dev_build_h_eff_noalloc.argtypes = [vp, vp, vp, P, M, C.c_int, vp]
    L.llgmi_dev_add_demag_field_pbc.argtypes = [vp, vp, vp, M, C.c_uint64, C.c_uint64]
    L.llgmi_dev_rk45_step.argtypes = [vp, vp, P, M, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, dp, ip, dp]
    L.llgmi_dev_rk45_step_batch.argtypes = [C.POINTER(vp), C.POINTER(vp), P, M, C.c_int64, C.c_int,
                                            C.c_double, C.c_double, C.c_double, C.c_double, dp, ip, dp]
    L.llgmi_dev_rk45_run_batch.argtypes = [C.POINTER(vp), C.POINTER(vp), P, M, C.c_int64, C.c_int64, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_double, dp, ip, dp]
    L.llgmi_dev_rk45_run.argtypes = [vp, vp, P, M, C.c_int64, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, dp, ip, dp]
    L.llgmi_dev_upload.argtypes = [vp, dp, vp]
    L.llgmi_dev_download.argtypes = [vp, vp, dp]
    L.llgmi_profile_enable.argtypes = [vp, C.c_int]
    L.llgmi_profile_reset.argtypes = [vp]
    L.llgmi_profile_read.argtypes = [vp, C.c_int, dp, C.POINTER(C.c_int64)]
    L.llgmi_profile_fft_passes.argtypes = [vp, vp, C.c_double, C.c_int, dp]
    L.llgmi_mg_config_default.argtypes = [C.POINTER(MGConfig)]
    L.llgmi_mg_config_from_env.argtypes = [C.POINTER(MGConfig)]
    L.llgmi_ctx_set_mg_config.argtypes = [vp, C.POINTER(MGConfig)]
    L.llgmi_add_demag_field_poisson_mg.argtypes = [vp, dp, dp, M]
    L.llgmi_compute_demag_field_poisson_mg.argtypes = [vp, dp, dp, M]
    L.llgmi_mg_last_stats.argtypes = [vp, C.POINTER(MGStats)]
    L.llgmi_mg_hybrid_config_default.argtypes = [C.POINTER(MGHybridConfig)]
    L.llgmi_mg_hybrid_config_from_env.argtypes = [C.POINTER(MGHybridConfig)]
    L.llgmi_ctx_set_mg_hybrid_config.argtypes = [vp, C.POINTER(MGHybridConfig)]
    RS, RR = C.POINTER(RelaxSettings), C.POINTER(RelaxReport)
    L.llgmi_relax_settings_default.argtypes = [RS]
    L.llgmi_relax_settings_mumax_like.argtypes = [RS]
    L.llgmi_rk23_relax_step.argtypes = step_args
    L.llgmi_dev_rk23_relax_step.argtypes = [vp, vp, P, M, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_double, dp, ip, dp]
    L.llgmi_rk23_invalidate.argtypes = [vp]
    L.llgmi_rk23_last_b_eff.argtypes = [vp, dp, ip]
    L.llgmi_compute_energy_geom.argtypes = [vp, dp, M, dp, u8p, C.POINTER(Energy)]
    L.llgmi_dev_compute_energy_geom.argtypes = [vp, vp, M, dp, vp, C.POINTER(Energy)]
    L.llgmi_relax_with_report.argtypes = [vp, dp, P, M, C.c_int, RS, RR]
    L.llgmi_dev_relax_with_report.argtypes = [vp, vp, P, M, C.c_int, RS, RR]
    OM = C.POINTER(OvfMeta)
    L.llgmi_write_ovf2_text.argtypes = [C.c_char_p, G, dp, OM]
    L.llgmi_write_ovf2_binary4.argtypes = [C.c_char_p, G, dp, OM]
    L.llgmi_dev_write_ovf2.argtypes = [vp, vp, C.c_char_p, OM, C.c_int]
    L.llgmi_ovf_total_sim_time_line.argtypes = [C.c_double, C.c_char_p, C.c_size_t]
    L.llgmi_demag_cache_path.argtypes = [G, C.c_uint64, C.c_uint64, C.c_char_p, C.c_size_t]
    L.llgmi_demag_cache_save.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_char_p]
    L.llgmi_demag_cache_load.argtypes = [vp, C.c_uint64, C.c_uint64, C.c_char_p, ip]
    L.llgmi_mg_box.argtypes = [vp, C.POINTER(C.c_int64)]
    L.llgmi_mg_phi_read.argtypes = [vp, vp]
    L.llgmi_mg_phi_write.argtypes = [vp, vp]
    fp = C.POINTER(C.c_float)
    L.llgmi_ws_create.argtypes = [C.POINTER(WSShape), C.c_uint64, C.c_int, C.POINTER(vp)]
    L.llgmi_ws_destroy.argtypes = [vp]
    L.llgmi_ws_param_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.llgmi_ws_get_params.argtypes = [vp, fp]
    L.llgmi_ws_set_params.argtypes = [vp, fp]
    L.llgmi_ws_get_grads.argtypes = [vp, fp]
    L.llgmi_ws_forward.argtypes = [vp, vp, vp]
    L.llgmi_ws_loss_grad.argtypes = [vp, vp, vp, dp]
    L.llgmi_ws_adam_step.argtypes = [vp, C.c_double]
    L.llgmi_ws_train_step.argtypes = [vp, vp, vp, C.c_double, dp]
    L.llgmi_ws_pack_input.argtypes = [vp, C.c_int64, vp, C.c_int, C.c_int64, C.c_int64, C.c_int64,
                                      C.c_int64, vp]
    L.llgmi_ws_pack_phi.argtypes = [vp, C.c_int64, vp, C.c_int64, C.c_int64, C.c_int64, C.c_double, vp]
    L.llgmi_ws_unpack_phi.argtypes = [vp, C.c_int64, vp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int,
                                      vp]
    L.llgmi_dev_mask_cells.argtypes = [vp, vp, vp]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != OK:
        raise LlgmiError(rc, lib().llgmi_last_error().decode())


def device_count() -> int:
    c = C.c_int(0)
    _check(lib().llgmi_device_count(C.byref(c)))
    return c.value


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(x) -> Optional[int]:
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


class DeviceArray:
    """Device memory from the HIP runtime the library runs on (hipMalloc
    through ctypes): buffers for the llgmi_dev_* / llgmi_ws_* entry points
    without a second runtime in the process.  Copies are synchronous."""

    _hip = None

    @classmethod
    def _rt(cls):
        if cls._hip is None:
            lib()                                  # loads libamdhip64 with the library
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        rc = self._rt().hipMalloc(C.byref(p), int(nbytes))
        if rc != 0:
            raise LlgmiError(f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr, self.nbytes = int(p.value), int(nbytes)

    def data_ptr(self) -> int:
        return self.ptr

    def at(self, offset: int) -> int:
        """Device address offset bytes into the buffer."""
        if not 0 <= offset <= self.nbytes:
            raise ValueError("offset outside the buffer")
        return self.ptr + int(offset)

    def _copy(self, dst: int, src: int, n: int, kind: int) -> None:
        rc = self._rt().hipMemcpy(dst, src, int(n), kind)
        if rc == 0 and kind != 2:
            # host-to-device copies from pageable memory (and device-to-device
            # copies) can return before the data has landed; the library's
            # streams do not wait for the null stream, so finish them here
            rc = self._rt().hipDeviceSynchronize()
        if rc != 0:
            raise LlgmiError(f"hipMemcpy failed: {rc}")

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload beyond the buffer")
        self._copy(self.ptr + int(offset), a.ctypes.data, a.nbytes, 1)

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(int(count), dtype=dtype)
        if offset + out.nbytes > self.nbytes:
            raise ValueError("download beyond the buffer")
        self._copy(out.ctypes.data, self.ptr + int(offset), out.nbytes, 2)
        return out

    def copy_from(self, src: "DeviceArray", nbytes: int, src_offset: int = 0, offset: int = 0) -> None:
        if src_offset + nbytes > src.nbytes or offset + nbytes > self.nbytes:
            raise ValueError("copy beyond a buffer")
        self._copy(self.ptr + int(offset), src.ptr + int(src_offset), nbytes, 3)

    def zero(self) -> None:
        rc = self._rt().hipMemset(self.ptr, 0, self.nbytes)
        if rc == 0:
            rc = self._rt().hipDeviceSynchronize()
        if rc != 0:
            raise LlgmiError(f"hipMemset failed: {rc}")

    def close(self) -> None:
        if self.ptr:
            self._rt().hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WarmStartNet:
    """The learned initial guess of the multigrid solve (llgmi_ws_*): a U-Net
    trained on the device.  Inputs x: bf16 [batch, H, W, 16]; targets: bf16
    [batch, H, W, c_pad]; outputs: FP32 [batch, H, W, c_pad] (device buffers,
    e.g. torch tensors of dtype bfloat16 / float32)."""

    def __init__(self, batch: int, height: int, width: int, out_channels: int, base_channels: int = 32,
                 in_channels: int = 4, valid_width: Optional[int] = None, valid_height: Optional[int] = None,
                 seed: int = 1, device: int = 0):
        self.shape = WSShape(batch, height, width, in_channels, out_channels, base_channels, 0,
                             valid_width or width, valid_height or height)
        self.c_pad = (out_channels + 15) // 16 * 16
        self._h = C.c_void_p()
        _check(lib().llgmi_ws_create(C.byref(self.shape), seed, device, C.byref(self._h)))
        n = C.c_int64()
        _check(lib().llgmi_ws_param_count(self._h, C.byref(n)))
        self.n_params = int(n.value)

    def close(self) -> None:
        if self._h:
            lib().llgmi_ws_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        h = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().llgmi_ws_get_params(self._h, h.ctypes.data_as(C.POINTER(C.c_float))))
        return h

    def set_params(self, p: np.ndarray) -> None:
        h = np.ascontiguousarray(p, dtype=np.float32)
        if h.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {h.size}")
        _check(lib().llgmi_ws_set_params(self._h, h.ctypes.data_as(C.POINTER(C.c_float))))

    def get_grads(self) -> np.ndarray:
        h = np.zeros(self.n_params, dtype=np.float32)
        _check(lib().llgmi_ws_get_grads(self._h, h.ctypes.data_as(C.POINTER(C.c_float))))
        return h

    def forward(self, x_dev, y_dev) -> None:
        _check(lib().llgmi_ws_forward(self._h, _ptr(x_dev), _ptr(y_dev)))

    def loss_grad(self, x_dev, t_dev) -> float:
        loss = C.c_double()
        _check(lib().llgmi_ws_loss_grad(self._h, _ptr(x_dev), _ptr(t_dev), C.byref(loss)))
        return loss.value

    def adam_step(self, lr: float) -> None:
        _check(lib().llgmi_ws_adam_step(self._h, lr))

    def train_step(self, x_dev, t_dev, lr: float) -> float:
        loss = C.c_double()
        _check(lib().llgmi_ws_train_step(self._h, _ptr(x_dev), _ptr(t_dev), lr, C.byref(loss)))
        return loss.value

    def pack_input(self, sample: int, m_dev, m_is_f32: bool, nx: int, ny: int, offx: int, offy: int,
                   x_dev) -> None:
        _check(lib().llgmi_ws_pack_input(self._h, sample, _ptr(m_dev), int(m_is_f32), nx, ny, offx, offy,
                                         _ptr(x_dev)))

    def pack_phi(self, sample: int, phi_dev, px: int, py: int, pz: int, scale: float, t_dev) -> None:
        _check(lib().llgmi_ws_pack_phi(self._h, sample, _ptr(phi_dev), px, py, pz, scale, _ptr(t_dev)))

    def unpack_phi(self, sample: int, y_dev, px: int, py: int, pz: int, scale: float, phi_dev,
                   accumulate: bool = False) -> None:
        _check(lib().llgmi_ws_unpack_phi(self._h, sample, _ptr(y_dev), px, py, pz, scale, int(accumulate),
                                         _ptr(phi_dev)))


def dev_rk45_step_batch(ctxs, m_devs, params, mats, mask: int, max_err: float, headroom: float,
                        dt_min: float, dt_max: float):
    """One adaptive RK45 attempt for each of len(ctxs) independent films
    (llgmi_dev_rk45_step_batch): film i is ctxs[i] with device state m_devs[i],
    params[i] (its dt is updated in place) and mats[i].  Returns the per-film
    (eps, accepted, dt_used) tuples of step_llg_rk45_recompute_field_masked_adaptive."""
    k = len(ctxs)
    if not (len(m_devs) == len(params) == len(mats) == k):
        raise ValueError("ctxs, m_devs, params and mats must have one entry per film")
    hs = (C.c_void_p * max(k, 1))(*[c._h for c in ctxs])
    ms = (C.c_void_p * max(k, 1))(*[_ptr(m) for m in m_devs])
    ps = (_Params * max(k, 1))(*[p._c() for p in params])
    cms = (_Material * max(k, 1))(*[m._c() for m in mats])
    eps = np.zeros(max(k, 1))
    dtu = np.zeros(max(k, 1))
    acc = np.zeros(max(k, 1), dtype=np.intc)
    _check(lib().llgmi_dev_rk45_step_batch(hs, ms, ps, cms, k, mask, max_err, headroom, dt_min, dt_max,
                                           _dp(eps), acc.ctypes.data_as(C.POINTER(C.c_int)), _dp(dtu)))
    for i, p in enumerate(params):
        p.dt = ps[i].dt
    return [(float(eps[i]), bool(acc[i]), float(dtu[i])) for i in range(k)]


def dev_rk45_run_batch(ctxs, m_devs, params, mats, attempts: int, mask: int, max_err: float, headroom: float,
                       dt_min: float, dt_max: float):
    """`attempts` consecutive adaptive RK45 attempts of each film with the
    controller on the device (llgmi_dev_rk45_run_batch); params[i].dt is
    updated in place.  Returns out[a][i] = (eps, accepted, dt_used)."""
    k = len(ctxs)
    if not (len(m_devs) == len(params) == len(mats) == k):
        raise ValueError("ctxs, m_devs, params and mats must have one entry per film")
    hs = (C.c_void_p * max(k, 1))(*[c._h for c in ctxs])
    ms = (C.c_void_p * max(k, 1))(*[_ptr(m) for m in m_devs])
    ps = (_Params * max(k, 1))(*[p._c() for p in params])
    cms = (_Material * max(k, 1))(*[m._c() for m in mats])
    no = max(k * attempts, 1)
    eps = np.zeros(no)
    dtu = np.zeros(no)
    acc = np.zeros(no, dtype=np.intc)
    _check(lib().llgmi_dev_rk45_run_batch(hs, ms, ps, cms, k, attempts, mask, max_err, headroom, dt_min, dt_max,
                                          _dp(eps), acc.ctypes.data_as(C.POINTER(C.c_int)), _dp(dtu)))
    for i, p in enumerate(params):
        p.dt = ps[i].dt
    return [[(float(eps[a * k + i]), bool(acc[a * k + i]), float(dtu[a * k + i])) for i in range(k)]
            for a in range(attempts)]


class Context:
    """A grid + precision + GPU: the reference's operator caches and RK45Scratch."""

    def __init__(self, grid: Grid2D, precision: int = F64, device: int = 0):
        self.grid = grid
        self.precision = precision
        self._h = C.c_void_p()
        g = grid._c()
        _check(lib().llgmi_ctx_create(C.byref(g), precision, device, C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().llgmi_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n(self) -> int:
        return self.grid.nx * self.grid.ny

    def stream(self) -> int:
        s = C.c_void_p()
        _check(lib().llgmi_ctx_stream(self._h, C.byref(s)))
        return s.value or 0

    def synchronize(self) -> None:
        _check(lib().llgmi_ctx_synchronize(self._h))

    # ---- host (AoS f64) ----
    def build_h_eff_masked_geom(self, m, params: LLGParams, mat: Material, mask: int,
                                geom_mask=None) -> np.ndarray:
        m = _f64(m).reshape(-1, 3)
        b = np.zeros_like(m)
        p, cm = params._c(), mat._c()
        gm = None
        if geom_mask is not None:
            g8 = np.ascontiguousarray(np.asarray(geom_mask).reshape(-1).astype(np.uint8))
            gm = g8.ctypes.data_as(C.POINTER(C.c_uint8))
        _check(lib().llgmi_build_h_eff_masked_geom(self._h, _dp(m), _dp(b), C.byref(p),
                                                   C.byref(cm), mask, gm))
        return b

    def build_h_eff_masked(self, m, params: LLGParams, mat: Material, mask: int) -> np.ndarray:
        return self.build_h_eff_masked_geom(m, params, mat, mask, None)

    def build_h_eff(self, m, params: LLGParams, mat: Material) -> np.ndarray:
        return self.build_h_eff_masked_geom(m, params, mat, FieldMask.Full, None)

    def add_demag_field_pbc(self, m, b: np.ndarray, mat: Material, pbc_x: int = 0,
                            pbc_y: int = 0) -> None:
        m = _f64(m).reshape(-1, 3)
        assert b.dtype == np.float64 and b.flags.c_contiguous and b.size == m.size
        cm = mat._c()
        _check(lib().llgmi_add_demag_field_pbc(self._h, _dp(m), _dp(b), C.byref(cm), pbc_x, pbc_y))

    def add_demag_field(self, m, b: np.ndarray, mat: Material) -> None:
        self.add_demag_field_pbc(m, b, mat, 0, 0)

    def compute_demag_field(self, m, mat: Material) -> np.ndarray:
        m = _f64(m).reshape(-1, 3)
        b = np.zeros_like(m)
        self.add_demag_field_pbc(m, b, mat, 0, 0)
        return b

    # ---- multigrid Poisson demag (demag_poisson_mg.rs) ----
    def set_mg_config(self, cfg: Optional[MGConfig]) -> None:
        """Pins the MG configuration; None reads LLG_DEMAG_MG_* at every call."""
        _check(lib().llgmi_ctx_set_mg_config(self._h, C.byref(cfg) if cfg is not None else None))

    def set_mg_hybrid_config(self, cfg: Optional[MGHybridConfig]) -> None:
        """Pins the hybrid MG settings; None reads LLG_DEMAG_MG_HYBRID_* at every call."""
        _check(lib().llgmi_ctx_set_mg_hybrid_config(self._h, C.byref(cfg) if cfg is not None else None))

    def add_demag_field_poisson_mg(self, m, b: np.ndarray, mat: Material) -> None:
        m = _f64(m).reshape(-1, 3)
        assert b.dtype == np.float64 and b.flags.c_contiguous and b.size == m.size
        cm = mat._c()
        _check(lib().llgmi_add_demag_field_poisson_mg(self._h, _dp(m), _dp(b), C.byref(cm)))

    def compute_demag_field_poisson_mg(self, m, mat: Material) -> np.ndarray:
        m = _f64(m).reshape(-1, 3)
        b = np.zeros_like(m)
        cm = mat._c()
        _check(lib().llgmi_compute_demag_field_poisson_mg(self._h, _dp(m), _dp(b), C.byref(cm)))
        return b

    def mg_last_stats(self) -> dict:
        st = MGStats()
        _check(lib().llgmi_mg_last_stats(self._h, C.byref(st)))
        return st.as_dict()

    def mg_box(self) -> Tuple[int, int, int, int, int, int]:
        """(px, py, pz, offx, offy, offz) of the finest multigrid box."""
        d = (C.c_int64 * 6)()
        _check(lib().llgmi_mg_box(self._h, d))
        return tuple(int(v) for v in d)

    def mg_phi_read(self, phi_dev) -> None:
        _check(lib().llgmi_mg_phi_read(self._h, _ptr(phi_dev)))

    def mg_phi_write(self, phi_dev) -> None:
        _check(lib().llgmi_mg_phi_write(self._h, _ptr(phi_dev)))

    def step_llg_rk45_recompute_field_masked_adaptive(self, m: np.ndarray, params: LLGParams,
                                                      mat: Material, mask: int, max_err: float,
                                                      headroom: float, dt_min: float,
                                                      dt_max: float):
        """Advances m (float64, C-contiguous, 3n) in place; updates params.dt.
        Returns (eps, accepted, dt_used)."""
        assert m.dtype == np.float64 and m.flags.c_contiguous and m.size == 3 * self.n
        p, cm = params._c(), mat._c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        _check(lib().llgmi_rk45_step(self._h, _dp(m), C.byref(p), C.byref(cm), mask, max_err,
                                     headroom, dt_min, dt_max, C.byref(eps), C.byref(acc),
                                     C.byref(dtu)))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value

    def step_llg_rk45_recompute_field_adaptive(self, m, params, mat, max_err, headroom, dt_min,
                                               dt_max):
        return self.step_llg_rk45_recompute_field_masked_adaptive(
            m, params, mat, FieldMask.Full, max_err, headroom, dt_min, dt_max)

    def rk45_reset(self) -> None:
        _check(lib().llgmi_rk45_reset(self._h))

    # ---- relaxation ----
    def step_llg_rk23_recompute_field_masked_relax_adaptive(self, m: np.ndarray, params: LLGParams,
                                                            mat: Material, mask: int,
                                                            max_err: float, headroom: float,
                                                            dt_min: float, dt_max: float):
        """llg.rs:278-436: advances m (float64, C-contiguous) in place when accepted;
        updates params.dt; returns (eps, accepted, dt_used)."""
        assert m.dtype == np.float64 and m.flags.c_contiguous and m.size == 3 * self.n
        p, cm = params._c(), mat._c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        _check(lib().llgmi_rk23_relax_step(self._h, _dp(m), C.byref(p), C.byref(cm), mask,
                                           max_err, headroom, dt_min, dt_max, C.byref(eps),
                                           C.byref(acc), C.byref(dtu)))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value

    def rk23_invalidate_last_b_eff(self) -> None:
        _check(lib().llgmi_rk23_invalidate(self._h))

    def rk23_last_b_eff(self) -> Optional[np.ndarray]:
        b = np.zeros((self.n, 3))
        v = C.c_int()
        _check(lib().llgmi_rk23_last_b_eff(self._h, _dp(b), C.byref(v)))
        return b if v.value else None

    def compute_energy_geom(self, m, mat: Material, b_ext, geom=None) -> dict:
        """energy.rs:80-283 -> {exchange, anisotropy, zeeman, dmi, demag} (J)."""
        m = _f64(m).reshape(-1, 3)
        bx = _f64(b_ext).reshape(3)
        cm = mat._c()
        e = Energy()
        g = None if geom is None else np.ascontiguousarray(np.asarray(geom).astype(np.uint8).reshape(-1))
        if g is not None:
            assert g.size == self.n
        _check(lib().llgmi_compute_energy_geom(self._h, _dp(m), C.byref(cm), _dp(bx),
                                               None if g is None else g.ctypes.data_as(C.POINTER(C.c_uint8)),
                                               C.byref(e)))
        return e.as_dict()

    def compute_energy(self, m, mat: Material, b_ext) -> dict:
        return self.compute_energy_geom(m, mat, b_ext, None)

    def compute_total_energy(self, m, mat: Material, b_ext) -> float:
        e = self.compute_energy(m, mat, b_ext)
        return (((e["exchange"] + e["anisotropy"]) + e["zeeman"]) + e["dmi"]) + e["demag"]

    def relax_with_report(self, m: np.ndarray, params: LLGParams, mat: Material, mask: int,
                          settings: RelaxSettings) -> dict:
        """relax.rs:334-774: m and params.dt in place, settings.max_err written back."""
        assert m.dtype == np.float64 and m.flags.c_contiguous and m.size == 3 * self.n
        p, cm = params._c(), mat._c()
        rep = RelaxReport()
        _check(lib().llgmi_relax_with_report(self._h, _dp(m), C.byref(p), C.byref(cm), mask,
                                             C.byref(settings), C.byref(rep)))
        params.dt = p.dt
        return rep.as_dict()

    def dev_rk23_relax_step(self, m_dev, params: LLGParams, mat: Material, mask: int,
                            max_err: float, headroom: float, dt_min: float, dt_max: float):
        p, cm = params._c(), mat._c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        _check(lib().llgmi_dev_rk23_relax_step(self._h, _ptr(m_dev), C.byref(p), C.byref(cm), mask,
                                               max_err, headroom, dt_min, dt_max, C.byref(eps),
                                               C.byref(acc), C.byref(dtu)))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value

    def dev_compute_energy_geom(self, m_dev, mat: Material, b_ext, geom_dev=None) -> dict:
        bx = _f64(b_ext).reshape(3)
        cm = mat._c()
        e = Energy()
        _check(lib().llgmi_dev_compute_energy_geom(self._h, _ptr(m_dev), C.byref(cm), _dp(bx),
                                                   _ptr(geom_dev), C.byref(e)))
        return e.as_dict()

    def dev_relax_with_report(self, m_dev, params: LLGParams, mat: Material, mask: int,
                              settings: RelaxSettings) -> dict:
        p, cm = params._c(), mat._c()
        rep = RelaxReport()
        _check(lib().llgmi_dev_relax_with_report(self._h, _ptr(m_dev), C.byref(p), C.byref(cm),
                                                 mask, C.byref(settings), C.byref(rep)))
        params.dt = p.dt
        return rep.as_dict()

    # ---- on-disk formats ----
    def demag_cache_save(self, pbc_x: int = 0, pbc_y: int = 0, path: Optional[str] = None) -> None:
        """write_kernel_kspace (demag_fft_uniform.rs:282-320) of this context's operator."""
        _check(lib().llgmi_demag_cache_save(self._h, pbc_x, pbc_y, None if path is None else path.encode()))

    def demag_cache_load(self, pbc_x: int = 0, pbc_y: int = 0, path: Optional[str] = None) -> bool:
        """try_load_kernel_kspace (demag_fft_uniform.rs:322-429): True when loaded."""
        ok = C.c_int()
        _check(lib().llgmi_demag_cache_load(self._h, pbc_x, pbc_y, None if path is None else path.encode(),
                                            C.byref(ok)))
        return bool(ok.value)

    def dev_write_ovf2(self, m_dev, path: str, meta: Optional[OvfMeta] = None, binary4: bool = True) -> None:
        _check(lib().llgmi_dev_write_ovf2(self._h, _ptr(m_dev), path.encode(),
                                          None if meta is None else C.byref(meta), int(binary4)))

    # ---- device (SoA, context precision) ----
    def dev_build_h_eff(self, m_dev, b_dev, params: LLGParams, mat: Material, mask: int,
                        geom_dev=None) -> None:
        p, cm = params._c(), mat._c()
        _check(lib().llgmi_dev_build_h_eff(self._h, _ptr(m_dev), _ptr(b_dev), C.byref(p),
                                           C.byref(cm), mask, _ptr(geom_dev)))

    def dev_build_h_eff_noalloc(self, m_dev, b_dev, params: LLGParams, mat: Material, mask: int,
                                geom_dev=None) -> None:
        """dev_build_h_eff with no allocation (llgmi_dev_build_h_eff_noalloc):
        the scratch is the library's per-GPU static array; the context's FFT
        operator must already be set up by one dev_build_h_eff call."""
        p, cm = params._c(), mat._c()
        _check(lib().llgmi_dev_build_h_eff_noalloc(self._h, _ptr(m_dev), _ptr(b_dev), C.byref(p),
                                                   C.byref(cm), mask, _ptr(geom_dev)))

    def dev_add_demag_field_pbc(self, m_dev, b_dev, mat: Material, pbc_x=0, pbc_y=0) -> None:
        cm = mat._c()
        _check(lib().llgmi_dev_add_demag_field_pbc(self._h, _ptr(m_dev), _ptr(b_dev), C.byref(cm),
                                                   pbc_x, pbc_y))

    def dev_rk45_step(self, m_dev, params: LLGParams, mat: Material, mask: int, max_err: float,
                      headroom: float, dt_min: float, dt_max: float):
        p, cm = params._c(), mat._c()
        eps, dtu, acc = C.c_double(), C.c_double(), C.c_int()
        _check(lib().llgmi_dev_rk45_step(self._h, _ptr(m_dev), C.byref(p), C.byref(cm), mask,
                                         max_err, headroom, dt_min, dt_max, C.byref(eps),
                                         C.byref(acc), C.byref(dtu)))
        params.dt = p.dt
        return eps.value, bool(acc.value), dtu.value

    def dev_rk45_run(self, m_dev, params: LLGParams, mat: Material, mask: int, max_err: float, headroom: float,
                     dt_min: float, dt_max: float, attempts: int):
        """`attempts` consecutive adaptive RK45 attempts with the controller on
        the device (llgmi_dev_rk45_run); returns [(eps, accepted, dt_used)]."""
        p, cm = params._c(), mat._c()
        no = max(attempts, 1)
        eps, dtu = np.zeros(no), np.zeros(no)
        acc = np.zeros(no, dtype=np.intc)
        _check(lib().llgmi_dev_rk45_run(self._h, _ptr(m_dev), C.byref(p), C.byref(cm), attempts, mask, max_err,
                                        headroom, dt_min, dt_max, _dp(eps), acc.ctypes.data_as(C.POINTER(C.c_int)),
                                        _dp(dtu)))
        params.dt = p.dt
        return [(float(eps[a]), bool(acc[a]), float(dtu[a])) for a in range(attempts)]

    def dev_mask_cells(self, m_dev, keep_dev) -> None:
        """m = 0 where the keep bytes (device, one per cell) are 0."""
        _check(lib().llgmi_dev_mask_cells(self._h, _ptr(m_dev), _ptr(keep_dev)))

    def dev_upload(self, host_aos, dev_soa) -> None:
        h = _f64(host_aos).reshape(-1)
        _check(lib().llgmi_dev_upload(self._h, _dp(h), _ptr(dev_soa)))

    def dev_download(self, dev_soa) -> np.ndarray:
        h = np.zeros((self.n, 3))
        _check(lib().llgmi_dev_download(self._h, _ptr(dev_soa), _dp(h)))
        return h

    # ---- timing ----
    def profile_enable(self, classes=None) -> None:
        """Times the given kernel classes (iterable of K_*), all when None; [] stops."""
        if classes is None:
            mask = (1 << K_COUNT) - 1
        else:
            mask = 0
            for c in classes:
                mask |= 1 << int(c)
        _check(lib().llgmi_profile_enable(self._h, mask))

    def profile_reset(self) -> None:
        _check(lib().llgmi_profile_reset(self._h))

    def profile_read(self, cls: int):
        t, c = C.c_double(), C.c_int64()
        _check(lib().llgmi_profile_read(self._h, cls, C.byref(t), C.byref(c)))
        return t.value, c.value

    def profile_fft_passes(self, m_dev, ms: float, reps: int):
        """Average microseconds of the row R2C, column and row C2R passes, each
        launched reps times back to back (llgmi_profile_fft_passes)."""
        out = np.zeros(3)
        _check(lib().llgmi_profile_fft_passes(self._h, _ptr(m_dev), float(ms), int(reps), _dp(out)))
        return {"fft_row_fwd": float(out[0]), "fft_col_mul": float(out[1]), "fft_row_inv": float(out[2])}
