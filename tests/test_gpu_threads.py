"""The C ABI is safe to call from several host threads at once (the
reference's engine is reentrant and its plan cache thread-safe,
src/fpntt.jl:273-293): threads issuing single products, squares, short
products and host batches on their own buffers, with plans for new lengths
built while the others run, all get libgmp's limbs."""
import threading

import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import engine  # noqa: E402


def test_concurrent_calls():
    if not gmp_ref.available():
        pytest.skip("libgmp not found")
    engine.init()
    errors = []

    def worker(seed):
        try:
            rng = np.random.default_rng(seed)
            for it in range(12):
                m = int(rng.integers(160, 30000))
                n = int(rng.integers(152, m + 1))
                a, b = rand_big_style(rng, m, 1), rand_big_style(rng, n, 1)
                kind = (seed + it) % 4
                if kind == 0:
                    r = np.zeros(m + n, dtype=np.uint64)
                    engine.mul_fpntt2(r, 0, a, 0, m, b, 0, n)
                    ok = np.array_equal(r, gmp_ref.mul(a, b))
                elif kind == 1:
                    r = np.zeros(2 * m, dtype=np.uint64)
                    engine.sqr_fpntt2(r, 0, a, 0, m)
                    ok = np.array_equal(r, gmp_ref.sqr(a))
                elif kind == 2:
                    k = int(rng.integers(1, m + n))
                    r = np.zeros(k, dtype=np.uint64)
                    engine.mullo(r, 0, a, 0, m, b, 0, n, k)
                    ok = np.array_equal(r, gmp_ref.mul(a, b)[:k])
                else:
                    r = np.zeros(m + n + 2 * n, dtype=np.uint64)
                    engine.mul_batch([(r, 0, a, 0, m, b, 0, n, 0), (r, m + n, b, 0, n, None, 0, 0, 0)])
                    ok = np.array_equal(r[:m + n], gmp_ref.mul(a, b)) and np.array_equal(r[m + n:], gmp_ref.sqr(b))
                if not ok:
                    errors.append((seed, it, kind, m, n))
        except Exception as e:  # surfaced below
            errors.append((seed, repr(e)))

    threads = [threading.Thread(target=worker, args=(0x7A + s,)) for s in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


def test_single_products_across_streams():
    """Device-path single products share one workspace: calls issued on three
    different streams without host synchronisation, with host-path calls in
    between, must still each give their own product."""
    if not gmp_ref.available():
        pytest.skip("libgmp not found")
    from nbhip import device as D
    engine.init()
    rng = np.random.default_rng(0x5717)
    streams = [torch.cuda.Stream() for _ in range(3)]
    cases = []
    for i in range(24):
        m = int(rng.integers(200, 20000))
        n = int(rng.integers(152, m + 1))
        a, b = rand_big_style(rng, m, 1), rand_big_style(rng, n, 1)
        da, db = D.to_dev(a), D.to_dev(b)
        torch.cuda.synchronize()
        dr = D.empty_limbs(m + n)
        with torch.cuda.stream(streams[i % 3]):
            D.mul_dev(dr, da, db)
        if i % 5 == 4:                      # a host-path call on the library's stream in between
            r = np.zeros(m + n, dtype=np.uint64)
            engine.mul_fpntt2(r, 0, a, 0, m, b, 0, n)
            assert np.array_equal(r, gmp_ref.mul(a, b))
        cases.append((dr, a, b))
    torch.cuda.synchronize()
    for dr, a, b in cases:
        assert np.array_equal(D.to_host(dr), gmp_ref.mul(a, b))


def test_barrett_batches_created_and_destroyed_beside_products():
    """Barrett contexts and batches are created, run and destroyed on one
    thread while four others run host-path products: the destroy calls share
    the library's event pool and plan cache with every other call, so they
    must take the same lock (each result is still exact)."""
    if not gmp_ref.available():
        pytest.skip("libgmp not found")
    from nbhip import barrett as BR
    from nbhip import device as D
    from tests.helpers import ints, wl
    engine.init()
    errors = []
    stop = threading.Event()

    def products(seed):
        try:
            rng = np.random.default_rng(seed)
            while not stop.is_set():
                m = int(rng.integers(160, 12000))
                n = int(rng.integers(152, m + 1))
                a, b = rand_big_style(rng, m, 1), rand_big_style(rng, n, 1)
                r = np.zeros(m + n, dtype=np.uint64)
                engine.mul_fpntt2(r, 0, a, 0, m, b, 0, n)
                if not np.array_equal(r, gmp_ref.mul(a, b)):
                    errors.append(("product", seed, m, n))
        except Exception as e:  # surfaced below
            errors.append(("product", seed, repr(e)))

    def barrett():
        try:
            rng = np.random.default_rng(0xBA)
            for it in range(10):
                k = int(rng.integers(160, 2000))
                m = rng.integers(0, 2**64, k, dtype=np.uint64)
                m[-1] |= np.uint64(1)
                M = ints(m)
                ctx = BR.Barrett(D.to_dev(m))
                vals = [int.from_bytes(rng.bytes(16 * k), "little") for _ in range(4)]
                ts = [D.to_dev(wl.int_to_limbs(v, 2 * k)) for v in vals]
                rs = [D.empty_limbs(k) for _ in vals]
                RB = BR.BarrettBatch(ctx, list(zip(ts, rs)))
                RB.run()
                torch.cuda.synchronize()
                for r, v in zip(rs, vals):
                    if ints(D.to_host(r)) != v % M:
                        errors.append(("barrett", it, k))
                RB.close()
                ctx.close()
        except Exception as e:  # surfaced below
            errors.append(("barrett", repr(e)))

    workers = [threading.Thread(target=products, args=(0x51 + s,)) for s in range(4)]
    for t in workers:
        t.start()
    b = threading.Thread(target=barrett)
    b.start()
    b.join(timeout=100)
    stop.set()
    for t in workers:
        t.join(timeout=100)
    assert not b.is_alive() and not any(t.is_alive() for t in workers)
    assert not errors, errors
