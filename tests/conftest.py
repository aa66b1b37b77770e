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


class DeviceBuffer:
    """Device memory from the HIP runtime libllgmi.so is linked against
    (hipMalloc through ctypes), for the llgmi_dev_* entry points; freed on
    close() or garbage collection."""

    _hip = None

    def __init__(self, nbytes: int):
        import ctypes as C
        if DeviceBuffer._hip is None:
            import llgmi
            llgmi.lib()                       # loads libamdhip64.so.7 with the library
            DeviceBuffer._hip = C.CDLL("libamdhip64.so.7")
            DeviceBuffer._hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            DeviceBuffer._hip.hipFree.argtypes = [C.c_void_p]
        p = C.c_void_p()
        rc = DeviceBuffer._hip.hipMalloc(C.byref(p), nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value
        self.nbytes = nbytes

    def data_ptr(self) -> int:
        return self.ptr

    def upload(self, a, offset: int = 0) -> None:
        """Synchronous host-to-device copy of a numpy array at byte offset."""
        import ctypes as C
        import numpy as np
        a = np.ascontiguousarray(a)
        offset = int(offset)
        assert offset + a.nbytes <= self.nbytes
        DeviceBuffer._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rc = DeviceBuffer._hip.hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, 1)
        assert rc == 0, rc
        # a pageable upload can return before the DMA lands; the library's
        # streams do not wait for the null stream
        assert DeviceBuffer._hip.hipDeviceSynchronize() == 0

    def download(self, dtype, count: int, offset: int = 0):
        """Synchronous device-to-host copy of count elements of dtype."""
        import ctypes as C
        import numpy as np
        out = np.empty(int(count), dtype=dtype)
        offset = int(offset)
        assert offset + out.nbytes <= self.nbytes
        DeviceBuffer._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rc = DeviceBuffer._hip.hipMemcpy(out.ctypes.data, self.ptr + offset, out.nbytes, 2)
        assert rc == 0, rc
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: int, src_offset: int = 0, offset: int = 0) -> None:
        """Synchronous device-to-device copy."""
        import ctypes as C
        nbytes, src_offset, offset = int(nbytes), int(src_offset), int(offset)
        assert src_offset + nbytes <= src.nbytes and offset + nbytes <= self.nbytes
        DeviceBuffer._hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rc = DeviceBuffer._hip.hipMemcpy(self.ptr + offset, src.ptr + src_offset, nbytes, 3)
        assert rc == 0, rc
        assert DeviceBuffer._hip.hipDeviceSynchronize() == 0

    def close(self) -> None:
        if self.ptr:
            DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
