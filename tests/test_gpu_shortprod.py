"""Low short products on the device (mullo!/sqrlo!, src/mul.jl:166-378):
r[0:k] = a*b mod beta^k, bit-exact against the oracle's restatement and the
closed form, over the reference test's sweep (test/test_algorithms.jl:
134-200) and NTT-range sizes, through the host form, the device form and
batches mixing full and short products.  Exactly k limbs are written: the
limbs above stay as they were."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import ints, rand_big_style
from tests.test_shortprod_oracle import KS, KS_SQ, mullo_shapes, sqrlo_lens

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip._lib import NbhError  # noqa: E402

SENT = 0xDEADBEEFDEADBEEF
BETA = 1 << 64


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init()


def _rand(rng, n):
    return rng.integers(0, 2**64, n, dtype=np.uint64)


@pytest.mark.parametrize("k", KS)
def test_mullo_reference_sweep(k):
    rng = np.random.default_rng(0x10C4 + k)
    for la, lb in mullo_shapes(k):
        a, b = _rand(rng, la), _rand(rng, lb)
        want = O.mullo(a, b, k)
        for x, y in ((a, b), (b, a)):
            r = np.full(k + 2, SENT, dtype=np.uint64)
            engine.mullo(r, 0, x, 0, x.size, y, 0, y.size, k)
            assert np.array_equal(r[:k], want), (k, x.size, y.size)
            assert (r[k:] == SENT).all()


@pytest.mark.parametrize("k", KS_SQ)
def test_sqrlo_reference_sweep(k):
    rng = np.random.default_rng(0x50 + k)
    for la in sqrlo_lens(k):
        a = _rand(rng, la)
        r = np.full(k + 2, SENT, dtype=np.uint64)
        engine.sqrlo(r, 0, a, 0, la, k)
        assert np.array_equal(r[:k], O.sqrlo(a, k)), (k, la)
        assert (r[k:] == SENT).all()


@pytest.mark.parametrize("k", (O.MULLO_BASECASE_THRESHOLD - 1, O.MULLO_BASECASE_THRESHOLD + 1,
                               2 * O.MULLO_BASECASE_THRESHOLD + 3, 3000))
def test_mullo_all_ones(k):
    """Maximum carry chains across the truncation cut."""
    for la, lb in ((k, k), (k, k // 2 + 1)):
        a = np.full(la, 2**64 - 1, dtype=np.uint64)
        b = np.full(lb, 2**64 - 1, dtype=np.uint64)
        r = np.empty(k, dtype=np.uint64)
        engine.mullo(r, 0, a, 0, la, b, 0, lb, k)
        assert ints(r) == (ints(a) * ints(b)) % BETA**k


# NTT-range short products: (la, lb, k)
NTT_CASES = [(300, 300, 300), (1000, 1000, 1000), (4097, 2000, 4097), (2 * 6000, 6000, 6000),
             (20000, 20000, 20000), (700, 700, 2000), (100000, 100000, 100000), (65536, 40000, 70000)]


@pytest.mark.parametrize("la,lb,k", NTT_CASES)
def test_mullo_dev_against_the_full_product(la, lb, k):
    rng = np.random.default_rng(la + 7 * k)
    a, b = rand_big_style(rng, la, 1), rand_big_style(rng, lb, 1)
    da, db = D.to_dev(a), D.to_dev(b)
    r = D.to_dev(np.full(k + 2, SENT, dtype=np.uint64))
    D.mullo_dev(r, da, db, k)
    # the full product of the clamped operands, low k limbs
    ca, cb = D.to_dev(a[:min(la, k)]), D.to_dev(b[:min(lb, k)])
    full = D.empty_limbs(ca.numel() + cb.numel())
    D.mul_dev(full, ca, cb)
    torch.cuda.synchronize()
    got = D.to_host(r)
    ref = np.zeros(k, dtype=np.uint64)
    f = D.to_host(full)
    ref[:min(k, f.size)] = f[:k]
    assert np.array_equal(got[:k], ref)
    assert (got[k:] == SENT).all()
    if la * lb <= 1500 * 1500:
        assert ints(got[:k]) == (ints(a) * ints(b)) % BETA**k


@pytest.mark.parametrize("la,k", [(500, 500), (5000, 5000), (5000, 7000), (60000, 60000), (30000, 20000)])
def test_sqrlo_dev(la, k):
    rng = np.random.default_rng(la + k)
    a = rand_big_style(rng, la, 1)
    da = D.to_dev(a)
    r = D.to_dev(np.full(k + 1, SENT, dtype=np.uint64))
    D.sqrlo_dev(r, da, k)
    ca = D.to_dev(a[:min(la, k)])
    full = D.empty_limbs(2 * ca.numel())
    D.sqr_dev(full, ca)
    torch.cuda.synchronize()
    got = D.to_host(r)
    f = D.to_host(full)
    ref = np.zeros(k, dtype=np.uint64)
    ref[:min(k, f.size)] = f[:k]
    assert np.array_equal(got[:k], ref)
    assert got[k] == SENT
    if la <= 5000:
        assert ints(got[:k]) == (ints(a) ** 2) % BETA**k


def test_bits_of_the_whole_operand_bound_the_clamped_one():
    """bits_* describe the operands as passed; the library clamps them to k."""
    rng = np.random.default_rng(9)
    la, lb, k = 3000, 2500, 2000
    a, b = rand_big_style(rng, la, 1), rand_big_style(rng, lb, 1)
    r = D.empty_limbs(k)
    D.mullo_dev(r, D.to_dev(a), D.to_dev(b), k, 64 * la, 64 * lb)
    torch.cuda.synchronize()
    assert ints(D.to_host(r)) == (ints(a) * ints(b)) % BETA**k


def test_batch_mixes_full_and_short_products():
    """Short jobs (schoolbook and NTT, squares and products) beside full ones
    in one batch: each equals the product run alone; two runs and a graph
    replay agree."""
    rng = np.random.default_rng(21)
    specs = [(5000, 4000, 0), (5000, 4000, 4500), (300, None, 300), (300, 200, 0), (100, 80, 120),
             (20000, 20000, 20000), (150, None, 250), (64, 64, 64), (9000, None, 12000), (2000, 3000, 1)]
    jobs, want = [], []
    for la, lb, k in specs:
        a = rand_big_style(rng, la, 1 + len(jobs) % 2)
        b = None if lb is None else rand_big_style(rng, lb, 1)
        rl = k if k else la + (la if b is None else lb)
        r = D.to_dev(np.full(rl + 2, SENT, dtype=np.uint64))
        jobs.append((r, D.to_dev(a), None if b is None else D.to_dev(b), -1, -1, k))
        B = ints(a) if b is None else ints(b)
        v = ints(a) * B
        want.append((rl, v % BETA**k if k else v))
    batch = D.Batch(jobs)
    for rep in range(3):
        for j in jobs:
            j[0][:-2].fill_(-1)
        if rep < 2:
            batch.run()
        else:
            batch.replay()
        torch.cuda.synchronize()
        for j, (rl, v) in zip(jobs, want):
            got = D.to_host(j[0])
            assert ints(got[:rl]) == v
            assert (got[rl:] == SENT).all()
    batch.close()


def test_errors():
    a = D.to_dev(np.ones(10, dtype=np.uint64))
    r = D.empty_limbs(10)
    with pytest.raises(NbhError):
        D.mullo_dev(r, a, a, 0)
    with pytest.raises(NbhError):
        D.sqrlo_dev(r, a, -3)
    with pytest.raises(NbhError):
        D.Batch([(r, a, a, -1, -1, -1)])


@pytest.mark.parametrize("k", [300, 3000, 20000])
def test_limbs_at_or_above_k_are_not_read(k):
    """Operands longer than k with all-ones limbs above k: the chunk pack
    reads whole dwords past the clamped operand, which must count as zero."""
    rng = np.random.default_rng(k + 1)
    a = np.concatenate([_rand(rng, k), np.full(64, 2**64 - 1, dtype=np.uint64)])
    b = np.concatenate([_rand(rng, k - 5), np.full(70, 2**64 - 1, dtype=np.uint64)])
    for x, y in ((a, b), (b, a)):
        r = D.empty_limbs(k)
        D.mullo_dev(r, D.to_dev(x), D.to_dev(y), k)
        torch.cuda.synchronize()
        assert ints(D.to_host(r)) == (ints(x[:k]) * ints(y[:k])) % BETA**k
    r = D.empty_limbs(k)
    D.sqrlo_dev(r, D.to_dev(a), k)
    torch.cuda.synchronize()
    assert ints(D.to_host(r)) == (ints(a[:k]) ** 2) % BETA**k
