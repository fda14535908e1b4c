"""Continuous-batching server: requests admitted while others run give the
same limbs as the same products computed alone, and as libgmp."""
import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip import serve  # noqa: E402


def _requests(count, seed, hi=20000):
    jobs = wl.batch_jobs(count=count, lo=152, hi=hi, sqr_frac=0.3, max_ratio=8.0, workload="serve")
    out = []
    for k, j in enumerate(jobs):
        a = wl.operand("serve", seed * 10000 + k, 0, j.m)
        b = None if j.sqr else wl.operand("serve", seed * 10000 + k, 1, j.n)
        out.append((a, b))
    return out


def test_alone_equals_in_mix():
    engine.init()
    host = _requests(120, 1)
    srv = serve.Server(budget_bytes=96 << 20, nstreams=3, max_batch=24)
    reqs = []
    for i, (a, b) in enumerate(host):
        ta, tb = D.to_dev(a), None if b is None else D.to_dev(b)
        ba = engine.magnitude_bits(a, 0, a.size)
        bb = ba if b is None else engine.magnitude_bits(b, 0, b.size)
        reqs.append(srv.submit(serve.Request(ta, tb, ba, bb, keep=True)))
        if i % 7 == 0:
            srv.poll()              # later requests arrive while earlier ones run
    srv.drain()
    assert srv.batches >= 3
    assert len(srv.done) == len(host) and not srv.rejected
    for r, (a, b) in zip(reqs, host):
        got = D.to_host(r.out)
        want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
        assert np.array_equal(got, want), r.rid
        alone = D.empty_limbs(r.limbs)
        if b is None:
            D.sqr_dev(alone, r.a)
        else:
            D.mul_dev(alone, r.a, r.b)
        torch.cuda.synchronize()
        assert np.array_equal(D.to_host(alone), got)
    lat = serve.latency_stats(srv.done)
    assert lat["p99_ms"] >= lat["p50_ms"] > 0


def test_budget_rejects_oversized_request():
    engine.init()
    srv = serve.Server(budget_bytes=1 << 20, nstreams=1)
    a = wl.operand("serve", 99, 0, 40000)
    ba = engine.magnitude_bits(a, 0, a.size)
    big = srv.submit(serve.Request(D.to_dev(a), None, ba, ba))
    small_a = wl.operand("serve", 98, 0, 300)
    bs = engine.magnitude_bits(small_a, 0, 300)
    small = srv.submit(serve.Request(D.to_dev(small_a), None, bs, bs, keep=True))
    srv.drain()
    assert srv.rejected == [big]
    assert np.array_equal(D.to_host(small.out), gmp_ref.sqr(small_a))
