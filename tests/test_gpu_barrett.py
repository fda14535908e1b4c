"""Batched Barrett reduction and powermod on the device (barrett_reduce!,
src/barrett.jl:44-81; powermod_limbs' Barrett branch,
src/algorithms.jl:17-107) against Python big ints."""
import numpy as np
import pytest

from tests.helpers import ints, wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import barrett as BR  # noqa: E402
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

BETA = 1 << 64


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init()


def _modulus(rng, k, kind="rand"):
    if kind == "pow":                 # beta^(k-1): mu has k+2 limbs
        m = np.zeros(k, dtype=np.uint64)
        m[-1] = 1
        return m
    if kind == "ones":
        return np.full(k, 2**64 - 1, dtype=np.uint64)
    m = rng.integers(0, 2**64, k, dtype=np.uint64)
    if m[-1] == 0:
        m[-1] = 1
    return m


def _values(rng, M, k, count):
    top = BETA ** (2 * k)
    vals = [0, 1, M - 1, M, M + 1, top - 1, (M - 1) ** 2 % top, M * (M // 3) + M - 1]
    while len(vals) < count:
        vals.append(int.from_bytes(rng.bytes(16 * k), "little"))
    return [v % top for v in vals[:count]]


@pytest.mark.parametrize("k,kind", [(1, "rand"), (5, "rand"), (160, "rand"), (161, "pow"), (700, "ones"),
                                    (3001, "rand")])
def test_barrett_batch(k, kind):
    rng = np.random.default_rng(k)
    m = _modulus(rng, k, kind)
    M = ints(m)
    ctx = BR.Barrett(D.to_dev(m))
    assert ctx.lmu == (k + 2 if kind == "pow" else k + 1)
    count = 12
    ts = [D.empty_limbs(2 * k) for _ in range(count)]
    rs = [D.empty_limbs(k) for _ in range(count)]
    RB = BR.BarrettBatch(ctx, list(zip(ts, rs)))
    for rep in range(2):             # the planned batch re-runs on new contents
        vals = _values(rng, M, k, count)
        for t, v in zip(ts, vals):
            t.copy_(D.to_dev(wl.int_to_limbs(v, 2 * k)))
        for r in rs:
            r.fill_(-1)
        RB.run()
        torch.cuda.synchronize()
        for r, v in zip(rs, vals):
            assert ints(D.to_host(r)) == v % M, (k, kind, rep)
    RB.close()
    ctx.close()


@pytest.mark.parametrize("k,nbits,count", [(160, 3, 4), (200, 100, 6), (300, 30, 3), (1024, 70, 4)])
def test_powermod_batch(k, nbits, count):
    rng = np.random.default_rng(1000 + k)
    m = _modulus(rng, k)
    M = ints(m)
    e = (1 << (nbits - 1)) | (int.from_bytes(rng.bytes(16), "little") % (1 << (nbits - 1)))
    bases = []
    for i in range(count):
        b = int.from_bytes(rng.bytes(8 * k), "little") % M
        b = max(b, 2)
        lb = max(1, (b.bit_length() + 63) // 64)
        bases.append(wl.int_to_limbs(b, lb))
    P = BR.PowermodBatch(m, bases, e)
    out = P.run()
    torch.cuda.synchronize()
    for b, r in zip(bases, out):
        assert ints(D.to_host(r)) == pow(ints(b), e, M)
    # a second run of the planned exponentiation gives the same limbs
    first = [D.to_host(r) for r in out]
    out = P.run()
    torch.cuda.synchronize()
    assert all(np.array_equal(x, D.to_host(r)) for x, r in zip(first, out))
    P.close()


@pytest.mark.parametrize("k,nbits,count", [(200, 40, 3), (1024, 70, 4)])
def test_powermod_graph_replays(k, nbits, count):
    """The whole exponentiation captured as one graph: every replay, with
    eager device work and a cleared accumulator between replays, gives the
    eager run's limbs and pow()'s value."""
    rng = np.random.default_rng(2000 + k)
    m = _modulus(rng, k)
    M = ints(m)
    e = (1 << (nbits - 1)) | (int.from_bytes(rng.bytes(16), "little") % (1 << (nbits - 1)))
    bases = [wl.int_to_limbs(max(2, int.from_bytes(rng.bytes(8 * k), "little") % M), k) for _ in range(count)]
    P = BR.PowermodBatch(m, bases, e)
    P.run()
    torch.cuda.synchronize()
    eager = [D.to_host(r) for r in P.acc]
    for b, r in zip(bases, eager):
        assert ints(r) == pow(ints(b), e, M)
    P.capture()
    for rep in range(3):
        for r in P.acc:
            r.fill_(-1)
        x = torch.arange(1 << 16, device="cuda", dtype=torch.int64).sum()   # eager work between replays
        out = P.replay()
        torch.cuda.synchronize()
        assert int(x) == (1 << 16) * ((1 << 16) - 1) // 2
        for want, r in zip(eager, out):
            assert np.array_equal(D.to_host(r), want), rep
    P.close()


def test_barrett_arguments_are_checked():
    import ctypes
    from nbhip._lib import NBH_ERR_ARG, lib
    m = D.to_dev(np.array([5, 7], dtype=np.uint64))
    h = ctypes.c_void_p()
    # lmu above k + 2 and a zero k are refused
    assert lib().nbh_barrett_create(D._p(m), 2, D._p(m), 5, None, ctypes.byref(h)) == NBH_ERR_ARG
    assert lib().nbh_barrett_create(D._p(m), 0, D._p(m), 1, None, ctypes.byref(h)) == NBH_ERR_ARG
    # a spectrum needs a positive bound on the other operand
    assert lib().nbh_spectrum_create(D._p(m), 2, -1, 0, None, ctypes.byref(h)) == NBH_ERR_ARG
    assert lib().nbh_spectrum_info(None, None, None, None, None) == NBH_ERR_ARG
