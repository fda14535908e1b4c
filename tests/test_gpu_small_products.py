"""Products whose transform fits one LDS tile (N <= 8192: the whole transform
pair, pointwise product and inverse in the mid kernel, then the Garner unpack
and carries) and their neighbours just past it, through every entry point:
single products and squares on host buffers, short products and zero-top /
zero operands on device buffers, and batches mixing them with multi-pass
products (alone = in the mix, graph replays).  Checked against the oracle
(the reference's pipeline restated, src/fpntt.jl:982-1037 with the unpack of
:806-925) and libgmp, over the reference's operand styles (all-ones operands
give the longest carry chains)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gmp_ref
from tests.helpers import rand_big_style, wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

# (m, n): balanced, unbalanced and squares at transform lengths 512 .. 16384
SHAPES = [(152, 152), (160, 0), (200, 64), (257, 199), (300, 300), (511, 0), (700, 400), (1000, 64),
          (1400, 1400), (2000, 0), (2600, 2000), (3000, 1500), (5000, 400), (5600, 0), (5700, 5700)]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()
    engine.init()


def _want(a, b):
    return O.sqr(a) if b is None else O.mul(a, b)


@pytest.mark.parametrize("style", [1, 2, 3, 4])
def test_single_products_one_tile_range(style):
    rng = np.random.default_rng(100 + style)
    for m, n in SHAPES:
        a = rand_big_style(rng, m, style)
        b = None if n == 0 else rand_big_style(rng, n, style)
        N = engine.fp_ntt_params(engine.magnitude_bits(a, 0, m),
                                 engine.magnitude_bits(a if b is None else b, 0, m if b is None else n))[1]
        r = np.zeros(2 * m if b is None else m + n, dtype=np.uint64)
        if b is None:
            engine.sqr_fpntt2(r, 0, a, 0, m)
        else:
            engine.mul_fpntt2(r, 0, a, 0, m, b, 0, n)
        want = _want(a, b)
        assert np.array_equal(r, want), (m, n, style, N)
        assert np.array_equal(want, gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b))


def test_device_path_short_products_and_zero_tops():
    rng = np.random.default_rng(7)
    for m, n, k in [(300, 250, 1), (300, 250, 17), (300, 250, 549), (300, 250, 550), (2600, 2000, 2500),
                    (900, 0, 1000), (900, 0, 1800)]:
        a = rand_big_style(rng, m, 1)
        b = None if n == 0 else rand_big_style(rng, n, 1)
        r = torch.full((k + 2,), -1, dtype=torch.int64, device="cuda")
        if b is None:
            D.sqrlo_dev(r, D.to_dev(a), k)
        else:
            D.mullo_dev(r, D.to_dev(a), D.to_dev(b), k)
        torch.cuda.synchronize()
        got = D.to_host(r)
        full = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
        want = np.zeros(k, dtype=np.uint64)
        want[:min(k, full.size)] = full[:k]
        assert np.array_equal(got[:k], want), (m, n, k)
        assert (got[k:] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    # zero top limbs (magnitude below the buffer) and a zero operand
    a = rand_big_style(rng, 800, 1)
    a[-40:] = 0
    b = rand_big_style(rng, 600, 1)
    r = D.empty_limbs(1400)
    D.mul_dev(r, D.to_dev(a), D.to_dev(b))
    torch.cuda.synchronize()
    assert np.array_equal(D.to_host(r), gmp_ref.mul(a, b))
    z = np.zeros(500, dtype=np.uint64)
    r = torch.full((1100,), -1, dtype=torch.int64, device="cuda")
    D.mul_dev(r, D.to_dev(b), D.to_dev(z))
    torch.cuda.synchronize()
    assert not D.to_host(r).any()


def test_batch_small_and_multipass_jobs():
    """Small jobs (one mid tile per transform) beside larger ones (strided
    passes) in one planned batch: every product equal to libgmp and to the
    same product alone; a replay gives the same limbs."""
    rng = np.random.default_rng(0xF0)
    shapes = [(int(m), int(n)) for m, n in zip(rng.integers(152, 6000, 150), rng.integers(0, 6000, 150))]
    shapes += [(20000, 9000), (40000, 0), (3000, 12)]
    ops, tens = [], []
    for m, n in shapes:
        n = min(n, m)
        a = rand_big_style(rng, m, int(rng.integers(1, 5)))
        b = None if n == 0 else rand_big_style(rng, n, int(rng.integers(1, 5)))
        ops.append((a, b))
        tens.append((D.empty_limbs(m + (m if b is None else n)), D.to_dev(a), None if b is None else D.to_dev(b)))
    want = gmp_ref.products(ops)
    B = D.Batch(tens)
    try:
        for rep in range(2):
            for r, _, _ in tens:
                r.fill_(-1)
            B.run()
            torch.cuda.synchronize()
            for (r, _, _), w in zip(tens, want):
                assert np.array_equal(D.to_host(r), w)
    finally:
        B.close()
    for (r, ta, tb), w in list(zip(tens, want))[:40]:
        alone = D.empty_limbs(r.numel())
        if tb is None:
            D.sqr_dev(alone, ta)
        else:
            D.mul_dev(alone, ta, tb)
        torch.cuda.synchronize()
        assert np.array_equal(D.to_host(alone), w)


def test_batch_small_only_capture_replays():
    """A batch of small jobs only, captured as a graph and replayed with other
    device work between replays."""
    jobs = wl.batch_jobs(count=200, lo=152, hi=2500, workload="test")
    ops = wl.batch_operands(jobs, workload="test")
    want = gmp_ref.products(ops)
    tens = [(D.empty_limbs(w.size), D.to_dev(a), None if b is None else D.to_dev(b)) for (a, b), w in zip(ops, want)]
    B = D.Batch(tens)
    try:
        B.capture()
        for rep in range(3):
            for r, _, _ in tens:
                r.fill_(rep)
            B.replay()
            torch.cuda.synchronize()
            for (r, _, _), w in zip(tens, want):
                assert np.array_equal(D.to_host(r), w)
    finally:
        B.close()
