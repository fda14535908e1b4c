"""Shared test helpers: big-int conversion and the reference's adversarial
operand styles (test/test_fpntt.jl:120-131, test/test_differential.jl:6-19)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "nativebigint.jl_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from nbhip import workloads as wl  # noqa: E402


def ints(a) -> int:
    return wl.limbs_to_int(a)


def rand_big_style(rng, n: int, style: int | None = None) -> np.ndarray:
    """n-limb magnitude in [2^(64(n-1)), 2^64n): uniform, all-ones, 2^k +- 1,
    or a single high bit."""
    if style is None:
        style = int(rng.integers(1, 5))
    if style == 1:
        a = rng.integers(0, 2**64, n, dtype=np.uint64)
        if a[-1] == 0:
            a[-1] = 1
        return a
    if style == 2:
        return np.full(n, 2**64 - 1, dtype=np.uint64)
    if style == 3:
        k = int(rng.integers(64 * (n - 1), 64 * n))
        v = (1 << k) + int(rng.integers(-1, 2))
        v = max(v, 1 << (64 * (n - 1)))
        return wl.int_to_limbs(v, n)
    return wl.int_to_limbs(1 << (64 * n - 1), n)
