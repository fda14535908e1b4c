"""Pins the oracle's low short products (mullo!/sqrlo!, src/mul.jl:166-378)
before the device is checked against them: every case is the closed form
(a*b) mod beta^k in Python integers, over the sweep of the reference's own
test (test/test_algorithms.jl:134-200): k across the basecase -> Mulders ->
full-product crossovers, balanced, unbalanced, clamped (operands longer than
k), zero-filled (product shorter than k) and near-full shapes, both argument
orders, all-ones operands."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import ints

BETA = 1 << 64
Tb, Ts = O.MULLO_BASECASE_THRESHOLD, O.SQRLO_BASECASE_THRESHOLD
Tf, Tfs = O.MULLO_FULL_THRESHOLD, O.SQRLO_FULL_THRESHOLD
KS = (1, 2, 3, 5, 8, Tb - 1, Tb, Tb + 1, 2 * Tb + 3, Tf - 1, Tf, Tf + 17)
KS_SQ = (1, 2, 3, 5, 8, Ts - 1, Ts, Ts + 1, 2 * Ts + 3, Tfs - 1, Tfs, Tfs + 17)


def mullo_shapes(k):
    """test/test_algorithms.jl:170-173."""
    return [(k, k), (k, max(1, k // 2)), (k, max(1, 2 * k // 3)), (max(1, k - 1), max(1, k - 1)),
            (k, 1), (k, 2), (max(1, k // 3), max(1, k // 3)), (k + 4, k + 4), (2 * k, max(1, k // 2))]


def sqrlo_lens(k):
    """test/test_algorithms.jl:193-194."""
    return sorted({k, max(1, k - 1), max(1, 2 * k // 3), max(1, (k + 2) // 2), max(1, k // 2),
                   max(1, k // 3), k + 4})


def _low(v, k):
    return v % (BETA ** k)


@pytest.mark.parametrize("k", KS)
def test_mullo_sweep(k):
    rng = np.random.default_rng(0x10C4 + k)
    for la, lb in mullo_shapes(k):
        a = rng.integers(0, 2**64, la, dtype=np.uint64)
        b = rng.integers(0, 2**64, lb, dtype=np.uint64)
        want = _low(ints(a) * ints(b), k)
        assert ints(O.mullo(a, b, k)) == want, (k, la, lb)
        assert ints(O.mullo(b, a, k)) == want, (k, lb, la)


@pytest.mark.parametrize("k", (Tb - 1, Tb + 1, 2 * Tb + 3))
def test_mullo_all_ones(k):
    for la, lb in ((k, k), (k, k // 2 + 1)):
        a = np.full(la, 2**64 - 1, dtype=np.uint64)
        b = np.full(lb, 2**64 - 1, dtype=np.uint64)
        assert ints(O.mullo(a, b, k)) == _low(ints(a) * ints(b), k)


@pytest.mark.parametrize("k", KS_SQ)
def test_sqrlo_sweep(k):
    rng = np.random.default_rng(0x50 + k)
    for la in sqrlo_lens(k):
        a = rng.integers(0, 2**64, la, dtype=np.uint64)
        assert ints(O.sqrlo(a, k)) == _low(ints(a) ** 2, k), (k, la)


def test_sweep_reaches_the_mulders_cross_terms():
    """The sweep's Mulders-band shapes take both cross-term branches: with
    either one dropped, the oracle disagrees with the closed form."""
    k = Tf - 1
    assert Tb <= k < Tf
    m = O.mullo_split(k)
    rng = np.random.default_rng(1)
    a = rng.integers(0, 2**64, k, dtype=np.uint64)
    b = rng.integers(0, 2**64, k, dtype=np.uint64)
    assert k > m  # both operands reach past the split
    lo = _low(ints(a[:m]) * ints(b[:m]), k)
    assert lo != _low(ints(a) * ints(b), k)
    assert ints(O.mullo(a, b, k)) == _low(ints(a) * ints(b), k)
