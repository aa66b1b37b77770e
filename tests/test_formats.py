"""On-disk formats written by the library's host code (no GPU needed):
OVF 2.0 text and binary4 (ovf.rs:54-262) and the demag kernel cache file
name (demag_fft_uniform.rs:262-280).

The expected OVF files below are written out line by line from the
reference's writeln! sequence, with the numbers as Rust's `{:.Ne}` prints
them (exponent without '+' or leading zeros: 2.0000000000000001e-9, 0e0);
the kernel cache contents are checked on the GPU (tests/test_gpu_demag.py).
"""
import os
import struct

import numpy as np
import pytest

import llgmi as L

GRID = L.Grid2D(2, 1, 1e-9, 2e-9, 3e-9)
M = np.array([[0.6, 0.0, 0.8], [-0.5, 1e-12, 0.0]])

TEXT = """# OOMMF OVF 2.0
# Segment count: 1
# Begin: Segment
# Begin: Header
# Title: m
# meshtype: rectangular
# meshunit: m
# xmin: 0.0000000000000000e0
# ymin: 0.0000000000000000e0
# zmin: 0.0000000000000000e0
# xmax: 2.0000000000000001e-9
# ymax: 2.0000000000000001e-9
# zmax: 3.0000000000000000e-9
# valuedim: 3
# valuelabels: m_x m_y m_z
# valueunits: 1 1 1
# Desc: Total simulation time:  1.5000000000000000e-10  s
# xbase: 5.0000000000000003e-10
# ybase: 1.0000000000000001e-9
# zbase: 1.5000000000000000e-9
# xnodes: 2
# ynodes: 1
# znodes: 1
# xstepsize: 1.0000000000000001e-9
# ystepsize: 2.0000000000000001e-9
# zstepsize: 3.0000000000000000e-9
# End: Header
# Begin: Data Text
6.0000000000e-1 0.0000000000e0 8.0000000000e-1
-5.0000000000e-1 1.0000000000e-12 0.0000000000e0
# End: Data Text
# End: Segment
"""

BIN_HEADER = """# OOMMF OVF 2.0
# Segment count: 1
# Begin: Segment
# Begin: Header
# Title: B_demag
# Desc: one
# Desc: two
# meshunit: m
# meshtype: rectangular
# xbase: 5.00000000000000031e-10
# ybase: 1.00000000000000006e-9
# zbase: 1.49999999999999999e-9
# xstepsize: 1.00000000000000006e-9
# ystepsize: 2.00000000000000012e-9
# zstepsize: 2.99999999999999998e-9
# xnodes: 2
# ynodes: 1
# znodes: 1
# xmin: 0
# xmax: 2.00000000000000012e-9
# ymin: 0
# ymax: 2.00000000000000012e-9
# zmin: 0
# zmax: 2.99999999999999998e-9
# valuedim: 3
# valuelabels: B_x B_y B_z
# valueunits: T T T
# End: Header
# Begin: Data Binary 4
"""


def test_ovf2_text_matches_reference_layout(tmp_path):
    p = tmp_path / "a" / "b" / "m.ovf"          # parent directories are created
    meta = L.ovf_meta(desc_lines=[L.ovf_total_sim_time_line(1.5e-10)])
    L.write_ovf2_rectangular_text(str(p), GRID, M, meta)
    assert p.read_text() == TEXT
    # default meta = OvfMeta::magnetization() without desc lines
    L.write_ovf2_rectangular_text(str(p), GRID, M)
    assert p.read_text() == TEXT.replace("# Desc: Total simulation time:  1.5000000000000000e-10  s\n", "")


def test_ovf2_binary4_layout(tmp_path):
    p = tmp_path / "b.ovf"
    meta = L.ovf_meta(title="B_demag", desc_lines=["one", "two"], valuelabels=("B_x", "B_y", "B_z"),
                      valueunits=("T", "T", "T"))
    L.write_ovf2_rectangular_binary4(str(p), GRID, M, meta)
    raw = p.read_bytes()
    head = BIN_HEADER.encode()
    assert raw.startswith(head)
    body = raw[len(head):]
    assert struct.unpack("<f", body[:4])[0] == 1234567.0
    vals = np.frombuffer(body[4:4 + 24], dtype="<f4")
    np.testing.assert_array_equal(vals, M.astype(np.float32).reshape(-1))
    assert body[28:] == b"\n# End: Data Binary 4\n# End: Segment\n# End: File\n"


def test_ovf2_large_field_round_trip(tmp_path):
    g = L.Grid2D(37, 21, 3e-9, 4e-9, 1e-9)
    m = np.random.default_rng(3).standard_normal((g.n_cells(), 3))
    L.write_ovf2_rectangular_text(str(tmp_path / "t.ovf"), g, m)
    lines = (tmp_path / "t.ovf").read_text().splitlines()
    i0 = lines.index("# Begin: Data Text") + 1
    data = np.array([[float(x) for x in ln.split()] for ln in lines[i0:i0 + g.n_cells()]])
    np.testing.assert_allclose(data, m, rtol=1e-10, atol=0)   # {:.10e}
    assert lines[i0 + g.n_cells()] == "# End: Data Text"
    L.write_ovf2_rectangular_binary4(str(tmp_path / "b.ovf"), g, m)
    raw = (tmp_path / "b.ovf").read_bytes()
    k = raw.index(b"# Begin: Data Binary 4\n") + len(b"# Begin: Data Binary 4\n")
    vals = np.frombuffer(raw[k + 4:k + 4 + 12 * g.n_cells()], dtype="<f4").reshape(-1, 3)
    np.testing.assert_array_equal(vals, m.astype(np.float32))


@pytest.mark.parametrize("t,line", [
    (0.0, "Total simulation time:  0.0000000000000000e0  s"),
    (1e-9, "Total simulation time:  1.0000000000000001e-9  s"),
    (2.5e-12, "Total simulation time:  2.4999999999999998e-12  s"),   # the exact binary value
    (3.0, "Total simulation time:  3.0000000000000000e0  s"),
    (1.5e20, "Total simulation time:  1.5000000000000000e20  s"),
])
def test_rust_exponent_formatting(t, line):
    assert L.ovf_total_sim_time_line(t) == line


def test_demag_cache_path_is_the_references():
    assert L.demag_cache_path(L.Grid2D(128, 64, 4e-9, 4e-9, 3e-9)) == (
        "out/demag_cache/demag_kernel_nx128_ny64_pbcx0_pbcy0_px256_py128_"
        "dx4.000e-9_dy4.000e-9_dz3.000e-9_acc10.00.bin")
    assert L.demag_cache_path(L.Grid2D(100, 30, 2.5e-9, 5e-9, 1e-9), 2, 0) == (
        "out/demag_cache/demag_kernel_nx100_ny30_pbcx2_pbcy0_px100_py60_"
        "dx2.500e-9_dy5.000e-9_dz1.000e-9_acc10.00.bin")


def test_writer_errors(tmp_path):
    with pytest.raises(L.LlgmiError):
        L.write_ovf2_rectangular_text(str(tmp_path / "missing_parent_is_a_file" / "x.ovf"),
                                      L.Grid2D(0, 1, 1e-9, 1e-9, 1e-9), np.zeros((0, 3)))
    (tmp_path / "f").write_text("x")
    with pytest.raises(L.LlgmiError):
        L.write_ovf2_rectangular_text(str(tmp_path / "f" / "x.ovf"), GRID, M)   # parent is a file
