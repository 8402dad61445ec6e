"""The BF16 matrix-core operand layouts the surrogate kernels rely on,
checked against a float64 host GEMM of the same bf16 inputs (bf16 products
are exact in fp32; fp32 accumulation of 16 or 32 terms stays within 1e-6)."""
import ctypes as C

import numpy as np
import pytest

from pibiti_amd import sph

pytestmark = pytest.mark.gpu


def to_bf16_bits(x):
    b = np.asarray(x, np.float32).view(np.uint32)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)  # round to nearest even
    return r


def from_bf16_bits(u):
    return (u.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("variant,m,k,n", [(0, 32, 16, 32), (1, 16, 32, 16)])
def test_mfma_layout(variant, m, k, n):
    rng = np.random.default_rng(variant)
    a = to_bf16_bits(rng.standard_normal((m, k)))
    b = to_bf16_bits(rng.standard_normal((k, n)))
    c = np.zeros((m, n), np.float32)
    L = sph.lib()
    st = L.sph_mfma_selftest(variant, a.ctypes.data_as(C.POINTER(C.c_uint16)),
                             b.ctypes.data_as(C.POINTER(C.c_uint16)),
                             c.ctypes.data_as(C.POINTER(C.c_float)))
    assert st == 0
    ref = from_bf16_bits(a).astype(np.float64) @ from_bf16_bits(b).astype(np.float64)
    np.testing.assert_allclose(c, ref, rtol=1e-5, atol=1e-5)
