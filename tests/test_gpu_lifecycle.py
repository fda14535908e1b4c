"""Handle lifetimes and schedule changes through the C ABI.

- nbh_shutdown refuses while any handle is live (handles point into the plan
  and twiddle caches it frees); once they are destroyed it succeeds and the
  library initialises again.  No stale handle is ever run.
- A serving queue whose waiting requests were sized under one schedule
  re-sizes them when the schedule knobs change before they are admitted.
- Barrett batches created and destroyed on one host thread while another
  thread runs products (both touch the library's shared event pool).
"""
import ctypes
import gc
import threading

import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style, wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import _lib  # noqa: E402
from nbhip import barrett as BR  # noqa: E402
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip import serve  # noqa: E402


@pytest.fixture(autouse=True)
def _init():
    engine.init()


def test_shutdown_refused_while_a_handle_is_live():
    a, b = wl.operand("test", 401, 0, 900), wl.operand("test", 401, 1, 700)
    r = D.empty_limbs(1600)
    B = D.Batch([(r, D.to_dev(a), D.to_dev(b))])
    rc = _lib.lib().nbh_shutdown()
    assert rc == _lib.NBH_ERR_ARG
    assert b"live handle" in _lib.lib().nbh_last_error_string()
    # the refused call changed nothing: the handle still runs
    B.run()
    torch.cuda.synchronize()
    assert np.array_equal(D.to_host(r), gmp_ref.mul(a, b))
    B.close()
    gc.collect()
    rc = _lib.lib().nbh_shutdown()
    if rc != _lib.NBH_OK:
        pytest.skip("another test module still holds a handle: " + _lib.lib().nbh_last_error_string().decode())
    _lib._initialised = False
    engine.init()
    r2 = np.zeros(1600, dtype=np.uint64)
    engine.mul_fpntt2(r2, 0, a, 0, 900, b, 0, 700)
    assert np.array_equal(r2, gmp_ref.mul(a, b))


def test_server_resizes_waiting_requests_after_a_schedule_change():
    table = D.tuning_file_entries()
    srv = serve.NativeServer(budget_bytes=512 << 20, nstreams=1, max_batch=64)
    host, reqs = [], []
    try:
        for i in range(24):
            m = 3000 + 997 * i
            a, b = wl.operand("serve", 900 + i, 0, m), wl.operand("serve", 900 + i, 1, m - 100)
            req = serve.Request(D.to_dev(a), D.to_dev(b), engine.magnitude_bits(a, 0, m),
                                engine.magnitude_bits(b, 0, m - 100))
            req.out = D.empty_limbs(req.limbs)
            host.append((a, b))
            reqs.append(req)
        torch.cuda.synchronize()
        for req in reqs:
            srv.submit(req)                   # sized under the shipped schedule, none admitted yet
        # far smaller tiles: more tiles per transform, larger tile tables
        D.set_tuning_table([(1 << 20, (1024, 1024, 8, 64))])
        srv.drain()
        assert not srv.rejected and len(srv.done) == len(reqs)
        for req, (a, b) in zip(reqs, host):
            assert np.array_equal(D.to_host(req.out), gmp_ref.mul(a, b))
    finally:
        D.set_tuning_table(table)
        srv.close()


def test_barrett_batches_created_and_destroyed_beside_products():
    k = 300
    rng = np.random.default_rng(0xBA)
    m = rng.integers(0, 2**64, k, dtype=np.uint64)
    m[-1] |= np.uint64(1)
    M = wl.limbs_to_int(m)
    ctx = BR.Barrett(D.to_dev(m))
    errors = []

    def products():
        try:
            prng = np.random.default_rng(7)
            for _ in range(20):
                x, y = rand_big_style(prng, 2000, 1), rand_big_style(prng, 1500, 1)
                r = np.zeros(3500, dtype=np.uint64)
                engine.mul_fpntt2(r, 0, x, 0, 2000, y, 0, 1500)
                if not np.array_equal(r, gmp_ref.mul(x, y)):
                    errors.append("product")
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    def reductions():
        try:
            for it in range(10):
                vals = [int.from_bytes(rng.bytes(16 * k), "little") for _ in range(4)]
                ts = [D.to_dev(wl.int_to_limbs(v, 2 * k)) for v in vals]
                rs = [D.empty_limbs(k) for _ in vals]
                RB = BR.BarrettBatch(ctx, list(zip(ts, rs)))
                RB.run()
                torch.cuda.synchronize()
                for v, r in zip(vals, rs):
                    if wl.limbs_to_int(D.to_host(r)) != v % M:
                        errors.append(("reduction", it))
                RB.close()
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=products), threading.Thread(target=reductions)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    ctx.close()
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
