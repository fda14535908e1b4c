"""Products below the NTT thresholds inside batches run in the schoolbook
kernel (the reference's Karatsuba / mul_basecase! range, src/mul.jl:24-26,
127-164,383-389): bit-exact against libgmp, at the routing edges, with
all-ones operands (longest carry chains), beside NTT jobs in one batch, and
through the serving arena."""
import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style, wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip import serve  # noqa: E402

# (m, n) with n = 0 for a square; the routing edges of mul!/sqr!
SHAPES = [(1, 1), (1, 7), (7, 1), (2, 2), (28, 29), (63, 1000), (1000, 63), (64, 239), (239, 64),
          (151, 152), (152, 152), (3000, 63), (5000, 40), (0, 0)]
SQUARES = [1, 2, 3, 51, 52, 100, 159, 160]


def _want(a, b):
    return gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init()


def test_small_products_in_a_batch():
    rng = np.random.default_rng(21)
    host = []
    for k, (m, n) in enumerate(SHAPES):
        if m == 0:
            continue
        for style in (1, 2, 3):
            host.append((rand_big_style(rng, m, style), rand_big_style(rng, n, style)))
    for n in SQUARES:
        for style in (1, 2):
            host.append((rand_big_style(rng, n, style), None))
    # NTT-sized jobs in the same batch
    host.append((wl.operand("test", 1, 0, 3000), wl.operand("test", 1, 1, 2500)))
    host.append((wl.operand("test", 2, 0, 700), None))
    tens = []
    for a, b in host:
        r = D.empty_limbs(a.size + (a.size if b is None else b.size))
        tens.append((r, D.to_dev(a), None if b is None else D.to_dev(b)))
    B = D.Batch(tens)
    for rep in range(2):
        for r, _, _ in tens:
            r.fill_(-1)
        B.run()
        torch.cuda.synchronize()
        for (r, _, _), (a, b) in zip(tens, host):
            assert np.array_equal(D.to_host(r), _want(a, b)), (a.size, None if b is None else b.size, rep)
    B.close()


def test_zero_and_one_limb_values():
    zero = np.zeros(5, dtype=np.uint64)
    one = np.array([1], dtype=np.uint64)
    big = np.full(90, 2**64 - 1, dtype=np.uint64)
    jobs = [(zero, big), (big, zero), (one, big), (big, one), (zero, None)]
    tens = [(D.empty_limbs(a.size + (a.size if b is None else b.size)), D.to_dev(a),
             None if b is None else D.to_dev(b)) for a, b in jobs]
    B = D.Batch(tens)
    B.run()
    torch.cuda.synchronize()
    for (r, _, _), (a, b) in zip(tens, jobs):
        want = wl.int_to_limbs(wl.limbs_to_int(a) * wl.limbs_to_int(a if b is None else b), r.numel())
        assert np.array_equal(D.to_host(r), want)
    B.close()


def test_small_requests_served():
    rng = np.random.default_rng(5)
    srv = serve.Server(budget_bytes=64 << 20, nstreams=2, max_batch=16)
    reqs, host = [], []
    for i in range(60):
        m = int(rng.integers(1, 400))
        n = int(rng.integers(1, 200))
        a = rand_big_style(rng, m, 1 + i % 3)
        b = None if i % 5 == 0 else rand_big_style(rng, n, 1 + i % 3)
        ba = engine.magnitude_bits(a, 0, a.size)
        bb = ba if b is None else engine.magnitude_bits(b, 0, b.size)
        reqs.append(srv.submit(serve.Request(D.to_dev(a), None if b is None else D.to_dev(b), ba, bb, keep=True)))
        host.append((a, b))
        if i % 9 == 0:
            srv.poll()
    srv.drain()
    assert len(srv.done) == len(host)
    for r, (a, b) in zip(reqs, host):
        assert np.array_equal(D.to_host(r.out), _want(a, b))
