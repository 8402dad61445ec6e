"""The device square root used by the force kernel (sph_device.hpp,
sqrt_rn_normal) takes the hardware root s, which is within 1 ulp of the
true root, and moves it to a neighbour n = s -/+ 1 ulp when the residual
x - n * s of one fused multiply-add says n is nearer.  This pins the
correction on the CPU: for every starting point the hardware may return
(the correctly rounded root or either neighbour) it must end at numpy's
correctly rounded float32 square root.  The residual's sign is computed
exactly in float64 (n * s has at most 48 significant bits), as the fused
multiply-add's single rounding preserves it."""
import numpy as np


def corrected(x: np.ndarray, s: np.ndarray) -> np.ndarray:
    lo = (s.view(np.int32) - 1).view(np.float32)
    hi = (s.view(np.int32) + 1).view(np.float32)
    x64 = x.astype(np.float64)
    r = np.where(x64 - lo.astype(np.float64) * s.astype(np.float64) <= 0.0, lo, s)
    return np.where(x64 - hi.astype(np.float64) * s.astype(np.float64) > 0.0, hi, r)


def test_neighbour_correction_gives_the_rounded_root():
    rng = np.random.default_rng(7)
    x = np.concatenate([
        rng.uniform(1e-8, 1e-3, 400_000),                        # r^2 of SPH pairs
        np.exp(rng.uniform(np.log(2.0 ** -96), np.log(3e38), 400_000)),
        (rng.integers(1, 2 ** 23, 100_000) * 2.0 ** -40),         # exact squares and near-squares
    ]).astype(np.float32)
    squares = (np.arange(1, 50_000, dtype=np.float32) * np.float32(1.0 / 64)) ** 2
    x = np.concatenate([x, squares.astype(np.float32)])
    x = x[(x >= np.float32(2.0 ** -96)) & np.isfinite(x)]
    want = np.sqrt(x)
    for d in (-1, 0, 1):
        s = (want.view(np.int32) + d).view(np.float32)
        got = corrected(x, s)
        assert np.array_equal(got, want), d


def test_a_plain_hardware_root_would_not_be_exact():
    """The correction matters: one ulp off is a different float."""
    x = np.array([2.0, 3.0, 1e-6], np.float32)
    want = np.sqrt(x)
    off = (want.view(np.int32) + 1).view(np.float32)
    assert not np.array_equal(off, want)
    assert np.array_equal(corrected(x, off), want)
