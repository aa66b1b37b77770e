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
