"""Shared pytest setup: repo root on sys.path, the `gpu` marker, helpers."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


def has_gpu() -> bool:
    try:
        import torch  # noqa: F401  (device count without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def scaled_err(a, b, scale):
    """max_i |a_i - b_i| / scale_i over vector rows."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    if d.ndim == 2:
        d = d.max(axis=1)
    return float((d / np.maximum(scale, 1e-300)).max())


class DeviceArray:
    """Device memory from the HIP runtime the step library itself runs on,
    for tests that hand device pointers to the C ABI.  (torch may bring its
    own copy of the runtime: memory from a second runtime instance is not
    usable by the library's.)"""

    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C

            from pibiti_amd import sph
            sph.lib()  # load the library first: its runtime is the one bound below
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes: int):
        import ctypes as C
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        st = self.hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1))
        assert st == 0, f"hipMalloc failed: {st}"

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        assert cls.hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return d

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self) -> None:
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
