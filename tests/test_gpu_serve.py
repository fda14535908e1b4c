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


def _dev_requests(host, with_out=True):
    reqs = []
    for a, b in host:
        ta, tb = D.to_dev(a), None if b is None else D.to_dev(b)
        ba = engine.magnitude_bits(a, 0, a.size)
        bb = ba if b is None else engine.magnitude_bits(b, 0, b.size)
        r = serve.Request(ta, tb, ba, bb)
        if with_out:
            r.out = D.empty_limbs(r.limbs)
        reqs.append(r)
    return reqs


def test_native_server_alone_equals_in_mix():
    """The library's own serving queue (nbh_server_*): every ticket reported
    exactly once, results equal to libgmp and to the same product alone."""
    engine.init()
    host = _requests(120, 2)
    reqs = _dev_requests(host)
    srv = serve.NativeServer(budget_bytes=96 << 20, nstreams=3, max_batch=24, cap=16)
    try:
        for i, r in enumerate(reqs):
            srv.submit(r)
            if i % 7 == 0:
                srv.poll()          # later requests arrive while earlier ones run
        while not srv.idle():
            srv.poll()
        assert srv.batches >= 3
        assert not srv.rejected
        assert sorted(r.rid for r in srv.done) == list(range(len(host)))
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
    finally:
        srv.close()


def test_native_server_budget_drain_and_small_products():
    """A job larger than the budget is refused at submit; drain completes the
    rest; schoolbook-range products, zero operands and unnormalised top limbs
    (bit lengths read on the device) go through the same queue."""
    engine.init()
    srv = serve.NativeServer(budget_bytes=1 << 20, nstreams=2, max_batch=8)
    try:
        a = wl.operand("serve", 99, 0, 40000)
        ba = engine.magnitude_bits(a, 0, a.size)
        big = serve.Request(D.to_dev(a), None, ba, ba)
        big.out = D.empty_limbs(big.limbs)
        srv.submit(big)
        assert srv.rejected == [big]
        host = [(wl.operand("serve", 97, 0, 300), None), (wl.operand("serve", 96, 0, 20), wl.operand("serve", 96, 1, 9)),
                (np.zeros(200, dtype=np.uint64), wl.operand("serve", 95, 1, 180))]
        z = wl.operand("serve", 94, 0, 400)
        z[-3:] = 0                       # top limbs zero
        host.append((z, wl.operand("serve", 94, 1, 350)))
        reqs = _dev_requests(host)
        for r in reqs:
            r.bits_a = r.bits_b = 0      # read on the device
            srv.submit(r)
        srv.drain()
        assert srv.idle() and len(srv.done) == len(host)
        for r, (x, y) in zip(reqs, host):
            want = gmp_ref.sqr(x) if y is None else gmp_ref.mul(x, y)
            assert np.array_equal(D.to_host(r.out), want)
    finally:
        srv.close()


def test_native_memory_helpers():
    """nbh_alloc / nbh_upload / nbh_download / nbh_free round trip."""
    import ctypes
    from nbhip._lib import check, lib
    engine.init()
    x = np.arange(1000, dtype=np.uint64) * 0x9E3779B97F4A7C15
    p = ctypes.c_void_p()
    check(lib().nbh_alloc(8 * x.size, ctypes.byref(p)))
    try:
        check(lib().nbh_upload(p, x.ctypes.data, 8 * x.size))
        y = np.zeros_like(x)
        check(lib().nbh_download(y.ctypes.data, p, 8 * x.size))
        assert np.array_equal(x, y)
    finally:
        lib().nbh_free(p)


def test_native_server_short_products():
    """Low short products (nbh_job.k > 0) through the library's queue: exactly
    k limbs written, equal to libgmp's low limbs."""
    import ctypes
    from nbhip._lib import check, lib, nbh_job
    engine.init()
    srv = serve.NativeServer(budget_bytes=256 << 20, nstreams=2, max_batch=8)
    try:
        rng = np.random.default_rng(0x5407)
        cases = []
        for i in range(12):
            m = int(rng.integers(100, 9000))
            n = int(rng.integers(50, m + 1))
            k = int(rng.integers(1, m + n + 1))
            a, b = wl.operand("serve", 700 + i, 0, m), wl.operand("serve", 700 + i, 1, n)
            da, db = D.to_dev(a), D.to_dev(b)
            r = torch.full((k + 2,), -1, dtype=torch.int64, device="cuda")
            job = nbh_job(r.data_ptr(), da.data_ptr(), db.data_ptr(), m, n, 0, 0, k)
            t = ctypes.c_int64()
            check(lib().nbh_server_submit(srv.h, ctypes.byref(job), ctypes.byref(t)))
            cases.append((a, b, k, r, da, db))
        check(lib().nbh_server_drain(srv.h))
        for a, b, k, r, _, _ in cases:
            got = D.to_host(r)
            want = np.zeros(k, dtype=np.uint64)
            full = gmp_ref.mul(a, b)
            want[:min(k, full.size)] = full[:k]
            assert np.array_equal(got[:k], want), k
            assert (got[k:] == np.uint64(0xFFFFFFFFFFFFFFFF)).all()
    finally:
        srv.close()
