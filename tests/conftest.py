import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "llg-sim_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls libllgmi.so")


def rel_err(b_test, b_ref) -> float:
    """max_i |dB_i| / max_i |B_i| over cells (vector 2-norms)."""
    import numpy as np
    bt = np.asarray(b_test, dtype=np.float64).reshape(-1, 3)
    br = np.asarray(b_ref, dtype=np.float64).reshape(-1, 3)
    den = np.sqrt((br * br).sum(axis=1)).max()
    num = np.sqrt(((bt - br) ** 2).sum(axis=1)).max()
    if den == 0.0:
        return float(num)
    return float(num / den)


@pytest.fixture(scope="session")
def gpu_lib():
    import llgmi
    L = llgmi.lib()
    if llgmi.device_count() < 1:
        pytest.fail("no HIP device visible to libllgmi.so")
    return llgmi
