"""Runs the reference's own tests restated in C++ against the C++ host API
(tests/cpp/test_reference_api.cpp over llg-sim_amd/host/llg_sim.hpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_reference_api")


def test_reference_tests_through_cpp_host_api(gpu_lib):
    assert os.path.exists(BIN), "build with __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
