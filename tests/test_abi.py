"""C ABI checks that need no GPU: the library loads, exports every function
include/llgmi.h declares, and rejects bad arguments before touching a device."""
import ctypes as C
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(REPO, "include", "llgmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(llgmi_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    import llgmi
    assert _declared() == sorted(llgmi.EXPORTED)


def test_library_exports_every_declared_symbol():
    import llgmi
    L = C.CDLL(llgmi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(L, s)]
    assert not missing, missing


def test_struct_layouts_match_header():
    import llgmi
    assert C.sizeof(llgmi._Grid) == 40
    assert C.sizeof(llgmi._Material) == 72
    assert llgmi._Material.dmi.offset == 56
    assert C.sizeof(llgmi._Params) == 48


def test_invalid_arguments_rejected_without_device():
    import llgmi
    L = llgmi.lib()
    h = C.c_void_p()
    g = llgmi._Grid(0, 4, 1e-9, 1e-9, 1e-9)
    assert L.llgmi_ctx_create(C.byref(g), 0, 0, C.byref(h)) == llgmi.ERR_INVALID
    assert b"positive" in L.llgmi_last_error()
    g = llgmi._Grid(4, 4, -1.0, 1e-9, 1e-9)
    assert L.llgmi_ctx_create(C.byref(g), 0, 0, C.byref(h)) == llgmi.ERR_INVALID
    g = llgmi._Grid(4, 4, 1e-9, 1e-9, 1e-9)
    assert L.llgmi_ctx_create(C.byref(g), 5, 0, C.byref(h)) == llgmi.ERR_INVALID
    assert L.llgmi_build_h_eff(None, None, None, None, None) == llgmi.ERR_INVALID
    assert L.llgmi_ctx_destroy(None) == llgmi.OK


def test_context_creation_reports_missing_device():
    """With no GPU the call fails with a HIP error and a message (no abort);
    with a GPU it succeeds."""
    import llgmi
    if llgmi.device_count() > 0:
        ctx = llgmi.Context(llgmi.Grid2D(4, 4, 1e-9, 1e-9, 1e-9))
        ctx.close()
        return
    with pytest.raises(llgmi.LlgmiError) as e:
        llgmi.Context(llgmi.Grid2D(4, 4, 1e-9, 1e-9, 1e-9))
    assert e.value.code == llgmi.ERR_HIP


def test_mg_struct_layouts_and_defaults():
    import llgmi
    assert C.sizeof(llgmi.MGConfig) == 144
    assert C.sizeof(llgmi.MGStats) == 72
    c = llgmi.mg_config()
    # DemagPoissonMGConfig::default() (demag_poisson_mg.rs:166-185) and the
    # default operator iso27 / trilinear / Galerkin (:741-794)
    assert (c.pad_xy, c.n_vac_z, c.v_cycles, c.v_cycles_max) == (6, 16, 4, 64)
    assert (c.has_tol_abs, c.has_tol_rel, c.tol_rel) == (0, 1, 1e-6)
    assert (c.pre_smooth, c.post_smooth, c.smoother, c.warm_start) == (2, 2, 0, 1)
    assert c.omega == 2.0 / 3.0 and c.sor_omega == 1.0
    assert (c.bc, c.tree_theta, c.tree_leaf, c.tree_max_depth) == (2, 0.6, 64, 20)
    assert (c.stencil, c.prolong, c.coarse_op, c.iso27_alpha) == (2, 1, 1, 0.0)


_MG_ENV = [k for k in ("PAD_XY", "PAD_FACTOR_XY", "NVAC_Z", "VCYCLES", "VCYCLES_MAX", "TOL_ABS", "TOL_REL",
                       "PRE_SMOOTH", "POST_SMOOTH", "OMEGA", "SOR_OMEGA", "SMOOTHER", "WARM_START", "BC",
                       "TREE_THETA", "TREE_LEAF", "TREE_MAX_DEPTH", "STENCIL", "PROLONG", "COARSE_OP",
                       "ISO27_FLUX_ALPHA")]


def test_mg_config_from_env_follows_reference_parsing(monkeypatch):
    """DemagPoissonMGConfig::from_env (demag_poisson_mg.rs:188-267) and
    MGOperatorSettings::from_env (:741-794): aliases, clamps, fallbacks."""
    import llgmi
    for k in _MG_ENV:
        monkeypatch.delenv("LLG_DEMAG_MG_" + k, raising=False)
    d = llgmi.mg_config_from_env()
    assert bytes(d) == bytes(llgmi.mg_config())
    env = {"PAD_FACTOR_XY": " 0 ", "NVAC_Z": "3", "VCYCLES": "0", "VCYCLES_MAX": "200",
           "TOL_ABS": "-5", "TOL_REL": "3.5", "PRE_SMOOTH": "0", "POST_SMOOTH": "4",
           "OMEGA": "9", "SOR_OMEGA": "0.01", "SMOOTHER": " RBGS ", "WARM_START": "on",
           "BC": "Dipole", "TREE_THETA": "0.1", "TREE_LEAF": "2", "TREE_MAX_DEPTH": "100",
           "STENCIL": "7pt", "PROLONG": "injection", "COARSE_OP": "rd", "ISO27_FLUX_ALPHA": "-1"}
    for k, v in env.items():
        monkeypatch.setenv("LLG_DEMAG_MG_" + k, v)
    c = llgmi.mg_config_from_env()
    assert (c.pad_xy, c.n_vac_z, c.v_cycles, c.v_cycles_max) == (1, 3, 1, 200)
    assert (c.has_tol_abs, c.tol_abs, c.has_tol_rel, c.tol_rel) == (1, 0.0, 1, 1.0)
    assert (c.pre_smooth, c.post_smooth) == (1, 4)
    assert (c.omega, c.sor_omega) == (1.0, 0.2)
    assert (c.smoother, c.warm_start, c.bc) == (1, 1, 1)
    assert (c.tree_theta, c.tree_leaf, c.tree_max_depth) == (0.2, 8, 64)
    assert (c.stencil, c.prolong, c.coarse_op, c.iso27_alpha) == (0, 0, 0, 0.0)
    # PAD_XY wins over the legacy alias; unknown words fall back as the reference does
    monkeypatch.setenv("LLG_DEMAG_MG_PAD_XY", "9")
    monkeypatch.setenv("LLG_DEMAG_MG_SMOOTHER", "gauss")     # unknown: keeps the default
    monkeypatch.setenv("LLG_DEMAG_MG_WARM_START", "TRUE")    # case-sensitive match: off
    monkeypatch.setenv("LLG_DEMAG_MG_BC", "nonsense")        # unknown: keeps the default
    monkeypatch.setenv("LLG_DEMAG_MG_STENCIL", "13pt")       # unknown: iso27
    monkeypatch.setenv("LLG_DEMAG_MG_PROLONG", "cubic")      # unknown: trilinear
    monkeypatch.setenv("LLG_DEMAG_MG_COARSE_OP", "x")        # unknown: Galerkin
    monkeypatch.setenv("LLG_DEMAG_MG_ISO27_FLUX_ALPHA", "0.5")
    monkeypatch.setenv("LLG_DEMAG_MG_VCYCLES", "abc")         # unparsable: default
    c = llgmi.mg_config_from_env()
    assert (c.pad_xy, c.smoother, c.warm_start, c.bc) == (9, 0, 0, 2)
    assert (c.stencil, c.prolong, c.coarse_op, c.iso27_alpha, c.v_cycles) == (2, 1, 1, 0.5, 4)


def test_relax_struct_layouts_and_defaults():
    """llgmi_relax_settings / _report / llgmi_energy sizes, and the defaults of
    RelaxSettings::default() (relax.rs:107-139) and mumax_like_relax_settings
    (relax.rs:789-808) as the library fills them, against the oracle's."""
    import math
    import llgmi
    from oracle import oracle as O
    assert C.sizeof(llgmi.RelaxSettings) == 176
    assert C.sizeof(llgmi.RelaxReport) == 80
    assert C.sizeof(llgmi.Energy) == 40
    assert [f for f, _ in llgmi.RelaxSettings._fields_] == [f for f, _ in O.RelaxSettings._fields_]
    for lib_fn, orc_fn in ((llgmi.relax_settings, O.relax_settings),
                           (llgmi.mumax_like_relax_settings, O.mumax_like_relax_settings)):
        a, b = lib_fn(), orc_fn()
        for f, _ in llgmi.RelaxSettings._fields_:
            assert getattr(a, f) == getattr(b, f), f
    d = llgmi.relax_settings()
    assert (d.max_err, d.headroom, d.dt_min, d.dt_max) == (1e-5, 0.8, 1e-18, 1e-11)
    assert d.has_torque_threshold == 1 and d.torque_threshold == 1e-4
    assert d.torque_metric == llgmi.TORQUE_MAX and d.energy_stride == 3
    assert d.torque_plateau_checks == 0 and d.torque_plateau_min_checks == 5
    assert d.tighten_factor == math.sqrt(0.5) and d.tighten_floor == 1e-9   # FRAC_1_SQRT_2
    assert d.max_accepted_steps == 2_000_000 and d.has_final_torque_gate == 0
    mm = llgmi.mumax_like_relax_settings()
    assert mm.torque_metric == llgmi.TORQUE_MEANSQ and mm.has_torque_threshold == 0
    assert mm.rel_energy_tol == 0.0 and mm.torque_plateau_checks == 1
    assert mm.torque_plateau_min_checks == 1 and mm.torque_plateau_rel == 0.0


def test_relax_arguments_rejected_without_device():
    import llgmi
    L = llgmi.lib()
    s = llgmi.relax_settings(torque_metric=7)
    assert L.llgmi_relax_with_report(None, None, None, None, 2, C.byref(s), None) == llgmi.ERR_INVALID
    assert L.llgmi_rk23_relax_step(None, None, None, None, 2, 1e-5, 0.8, 1e-18, 1e-11,
                                   None, None, None) == llgmi.ERR_INVALID
    assert L.llgmi_relax_settings_default(None) == llgmi.ERR_INVALID


def test_mg_hybrid_config_from_env(monkeypatch):
    """HybridConfig::from_env (demag_poisson_mg.rs:296-356): off unless ENABLE is
    1/true/yes/on (trimmed, any case); then RADIUS, DELTA_VCYCLES (>= 1) and
    SIGMA (clamped to [0, 32]) are read."""
    import llgmi
    assert C.sizeof(llgmi.MGHybridConfig) == 48
    for k in ("ENABLE", "RADIUS", "DELTA_VCYCLES", "SIGMA"):
        monkeypatch.delenv("LLG_DEMAG_MG_HYBRID_" + k, raising=False)
    d = llgmi.mg_hybrid_config()
    assert (d.enabled, d.radius_xy, d.delta_v_cycles, d.sigma_cells, d.bmargin) == (0, 12, 60, 1.5, 2)
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_RADIUS", "5")
    assert llgmi.mg_hybrid_config_from_env().radius_xy == 12        # ignored while disabled
    for v, on in ((" On ", 1), ("YES", 1), ("true", 1), ("1", 1), ("0", 0), ("enable", 0)):
        monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_ENABLE", v)
        assert llgmi.mg_hybrid_config_from_env().enabled == on, v
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_ENABLE", "1")
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_DELTA_VCYCLES", "0")
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_SIGMA", "40")
    c = llgmi.mg_hybrid_config_from_env()
    assert (c.radius_xy, c.delta_v_cycles, c.sigma_cells) == (5, 1, 32.0)
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_SIGMA", "-3")
    assert llgmi.mg_hybrid_config_from_env().sigma_cells == 0.0
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_SIGMA", "NaN")
    assert llgmi.mg_hybrid_config_from_env().sigma_cells == 0.0
    monkeypatch.setenv("LLG_DEMAG_MG_HYBRID_SIGMA", "x")
    assert llgmi.mg_hybrid_config_from_env().sigma_cells == 1.5


def test_warm_start_shape_and_batch_arguments_rejected_without_device():
    """llgmi_ws_shape is 56 bytes; llgmi_ws_create rejects every bad shape
    and llgmi_dev_rk45_step_batch its bad arguments before any device call
    (so these run without a GPU); an empty batch is a no-op."""
    import llgmi
    L = llgmi.lib()
    assert C.sizeof(llgmi.WSShape) == 56
    h = C.c_void_p()
    good = dict(batch=1, height=64, width=64, in_channels=4, out_channels=36, base_channels=32,
                valid_width=60, valid_height=60)
    bad = [dict(height=60), dict(width=4), dict(batch=0), dict(in_channels=0), dict(in_channels=17),
           dict(out_channels=0), dict(out_channels=257), dict(base_channels=24), dict(base_channels=64),
           dict(valid_width=0), dict(batch=1 << 20, height=1024, width=1024)]
    for b in bad:
        kw = {**good, **b}
        s = llgmi.WSShape(kw["batch"], kw["height"], kw["width"], kw["in_channels"], kw["out_channels"],
                          kw["base_channels"], 0, kw["valid_width"], kw["valid_height"])
        assert L.llgmi_ws_create(C.byref(s), 1, 0, C.byref(h)) == 1, b
        assert not h.value
    assert L.llgmi_ws_create(None, 1, 0, C.byref(h)) == 1
    n = C.c_int64()
    assert L.llgmi_ws_param_count(None, C.byref(n)) == 1
    assert b"NULL" in L.llgmi_last_error()
    assert L.llgmi_ws_destroy(None) == 0
    # batched stepper: empty batch, negative count, NULL arrays
    z = C.c_double()
    assert L.llgmi_dev_rk45_step_batch(None, None, None, None, 0, 2, 1e-5, 0.8, 1e-18, 1e-11, None, None,
                                       None) == 0
    assert L.llgmi_dev_rk45_step_batch(None, None, None, None, -1, 2, 1e-5, 0.8, 1e-18, 1e-11, None, None,
                                       None) == 1
    assert L.llgmi_dev_rk45_step_batch(None, None, None, None, 3, 2, 1e-5, 0.8, 1e-18, 1e-11,
                                       C.byref(z), None, None) == 1
    assert L.llgmi_mg_box(None, None) == 1
    assert L.llgmi_dev_mask_cells(None, None, None) == 1
