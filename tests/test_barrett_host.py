"""Host logic of the device Barrett / powermod path, pinned by closed forms:
mu = floor(beta^(2k)/m) (barrett_mu!, src/barrett.jl:24-31), the sliding-
window schedule of powermod_limbs (src/algorithms.jl:17-107) against pow(),
and HAC 14.42's bound of two corrections that the correction kernel relies
on (barrett_reduce!'s while loop, src/barrett.jl:74-78)."""
import numpy as np
import pytest

from tests.helpers import wl
from nbhip.barrett import barrett_mu, exponent_schedule, window_bits

BETA = 1 << 64


def _rand_mod(rng, k):
    m = rng.integers(0, 2**64, k, dtype=np.uint64)
    if m[-1] == 0:
        m[-1] = 1
    return m


@pytest.mark.parametrize("k", [1, 2, 3, 17, 160])
def test_mu_closed_form(k):
    rng = np.random.default_rng(k)
    for _ in range(5):
        m = _rand_mod(rng, k)
        mu, lmu = barrett_mu(m)
        M, MU = wl.limbs_to_int(m), wl.limbs_to_int(mu)
        assert MU * M <= BETA ** (2 * k) < (MU + 1) * M
        assert lmu == k + 1 and mu[lmu - 1] != 0 and not mu[lmu:].any()
    if k >= 2:                       # m = beta^(k-1): mu = beta^(k+1), k+2 limbs
        m = np.zeros(k, dtype=np.uint64)
        m[-1] = 1
        mu, lmu = barrett_mu(m)
        assert lmu == k + 2 and wl.limbs_to_int(mu) == BETA ** (k + 1)


def test_mu_rejects_unnormalised_modulus():
    with pytest.raises(ValueError):
        barrett_mu(np.array([5, 0], dtype=np.uint64))


def _run_schedule(ops, b, m, w):
    table = [pow(b, 2 * j + 1, m) for j in range(1 << (w - 1))]
    acc = None
    for op, j in ops:
        if op == "load":
            acc = table[j]
        elif op == "sqr":
            acc = acc * acc % m
        else:
            acc = acc * table[j] % m
    return acc


def test_window_widths():
    assert [window_bits(n) for n in (1, 8, 9, 24, 25, 80, 81, 240, 241, 5000)] == [1, 1, 2, 2, 3, 3, 4, 4, 5, 5]


@pytest.mark.parametrize("nbits", [1, 2, 7, 8, 9, 24, 25, 80, 81, 240, 241, 700])
def test_schedule_computes_pow(nbits):
    rng = np.random.default_rng(nbits)
    for _ in range(20):
        e = (1 << (nbits - 1)) | int.from_bytes(rng.bytes(nbits // 8 + 1), "little") % (1 << (nbits - 1))
        m = int.from_bytes(rng.bytes(40), "little") | 1 | (1 << 319)
        b = int.from_bytes(rng.bytes(40), "little") % m
        ops = exponent_schedule(e)
        assert ops[0][0] == "load"
        assert _run_schedule(ops, b, m, window_bits(nbits)) == pow(b, e, m)
        # squarings: one per exponent bit below the leading window
        assert sum(op == "sqr" for op, _ in ops) <= nbits - 1


def _hac_corrections(T, M, MU, k):
    q1 = T >> (64 * (k - 1))
    q3 = (q1 * MU) >> (64 * (k + 1))
    r = (T - q3 * M) % BETA ** (k + 1)
    n = 0
    while r >= M:
        r -= M
        n += 1
    return r, n


@pytest.mark.parametrize("k", [1, 2, 5, 40])
def test_at_most_two_corrections(k):
    rng = np.random.default_rng(100 + k)
    mods = [_rand_mod(rng, k) for _ in range(4)]
    if k >= 2:
        e = np.zeros(k, dtype=np.uint64)
        e[-1] = 1
        mods.append(e)
    mods.append(np.full(k, 2**64 - 1, dtype=np.uint64))
    for m in mods:
        M = wl.limbs_to_int(m)
        MU = wl.limbs_to_int(barrett_mu(m)[0])
        Ts = [0, 1, M - 1, M, M + 1, BETA ** (2 * k) - 1, (M - 1) ** 2, M * M - 1 if M * M < BETA ** (2 * k) else 0]
        Ts += [int.from_bytes(rng.bytes(16 * k), "little") for _ in range(50)]
        for T in Ts:
            r, n = _hac_corrections(T, M, MU, k)
            assert r == T % M and n <= 2
