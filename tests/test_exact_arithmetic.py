"""Pins of two arithmetic identities the GPU kernels rely on for bit-exact
parity (CPU only):

* float division by a constant d equals rounding the double product x*(1/d) to
  float, for every float x (the kernels divide by the cell size, 6, 9, the
  diffusion coefficient and small counts this way);
* the magic-number integer division used to decode flat cell indices equals
  floor division for every index a grid can produce.
"""
import numpy as np
import pytest


def exact_division_mismatches(d: np.float32, x: np.ndarray) -> int:
    inv = 1.0 / np.float64(d)
    with np.errstate(all="ignore"):
        ieee = x / d
        fast = (x.astype(np.float64) * inv).astype(np.float32)
    ok = (ieee == fast) | (np.isnan(ieee) & np.isnan(fast))
    return int((~ok).sum())


@pytest.mark.parametrize("n", [3, 16, 24, 50, 64, 128, 256, 1000])
def test_division_by_constant_is_correctly_rounded(n):
    rng = np.random.default_rng(n)
    h = np.float32(10.0) / np.float32(n)
    lam = np.float32(1e-3) * np.float32(0.005) / (h * h)
    divisors = [h, np.float32(6), np.float32(9), np.float32(1) + np.float32(2) * lam,
                np.float32(3), np.float32(5)]
    for d in divisors:
        bits = rng.integers(0, 2**32, 400_000, dtype=np.uint64).astype(np.uint32)
        x = np.concatenate([
            (rng.random(400_000) * 2 - 1).astype(np.float32) * d,          # |x| < d (hat arguments)
            (rng.random(200_000) * 10).astype(np.float32),                  # positions in the domain
            bits.view(np.float32),                                          # every magnitude
            np.arange(-4096, 4096, dtype=np.float32) * d,                   # near multiples of d
            np.nextafter(np.arange(-4096, 4096, dtype=np.float32) * d, np.float32(np.inf)),
        ])
        x = x[np.isfinite(x)]
        assert exact_division_mismatches(d, x) == 0, d


def make_fastdiv(d: int):
    s = 0
    while (1 << s) < d:
        s += 1
    m = ((1 << 32) * ((1 << s) - d)) // d + 1
    return m & 0xFFFFFFFF, s


def fast_div(n: np.ndarray, m: int, s: int) -> np.ndarray:
    t = (n.astype(np.uint64) * np.uint64(m)) >> np.uint64(32)
    return ((t + n.astype(np.uint64)) >> np.uint64(s)).astype(np.int64)


@pytest.mark.parametrize("d", [1, 2, 3, 7, 16, 17, 129, 128 * 128, 129 * 129, 257 * 257, 1025 * 1025])
def test_fast_integer_division(d):
    m, s = make_fastdiv(d)
    rng = np.random.default_rng(d)
    n = np.concatenate([np.arange(0, 200_000), rng.integers(0, 2**31, 500_000),
                        np.array([2**31 - 1, 2**31 - 2, d - 1, d, d + 1, 2 * d - 1, 2 * d])])
    n = n[n < 2**31].astype(np.uint64)
    assert np.array_equal(fast_div(n, m, s), (n // np.uint64(d)).astype(np.int64))
