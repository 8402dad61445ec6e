"""sph_radix_sort_noalloc, the sort without any device allocation (its
scratch is a fixed device array), against sph_radix_sort, the entry point
that allocates its scratch per call, and numpy's stable sort: two calls in a
row, the second on a smaller input, so that the fixed array's left-over
contents from the first call must not matter.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import DeviceArray
from pibiti_amd import sph

pytestmark = pytest.mark.gpu

MAX = 1 << 20


def sort_with(fn, keys, vals, bits):
    host = np.empty(2 * len(keys), np.uint32)
    host[0::2] = keys
    host[1::2] = vals
    d = DeviceArray.from_numpy(host)
    t = DeviceArray(host.nbytes)
    st = fn(d.ptr, t.ptr, len(keys), bits, None)
    assert st == sph.SPH_OK, sph.strerror(st)
    out = d.to_numpy(np.uint32, host.size)
    d.free()
    t.free()
    return out[0::2].copy(), out[1::2].copy()


def check(n, bits, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    print("case: noalloc sort", n, bits, flush=True)
    k0, v0 = sort_with(sph.lib().sph_radix_sort_noalloc, keys, vals, bits)
    k1, v1 = sort_with(sph.lib().sph_radix_sort, keys, vals, bits)
    assert np.array_equal(k0, k1) and np.array_equal(v0, v1)
    order = np.argsort(keys & np.uint32((1 << bits) - 1) if bits < 32 else keys, kind="stable")
    assert np.array_equal(k0, keys[order]) and np.array_equal(v0, vals[order])


def test_noalloc_matches_the_allocating_sort_twice_in_a_row():
    check(MAX, 19, 1)        # the largest size the fixed array serves, 3 passes
    check(123_457, 16, 2)    # smaller, 2 passes: fewer tiles than the first call
    check(5_000, 8, 3)       # one pass


def test_noalloc_rejects_sizes_above_its_maximum():
    f = sph.lib().sph_radix_sort_noalloc
    assert f(C.c_void_p(8), C.c_void_p(16), MAX + 1, 16, None) == 8  # SPH_ERR_CAPACITY
    assert f(C.c_void_p(8), C.c_void_p(16), 0, 16, None) == sph.SPH_OK
