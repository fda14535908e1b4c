"""The benchmark configurations at their full sizes, through the C ABI.

- batch: the 4096-product list (limbs log-uniform in [152, 65536], 25%
  squares, ratio U[1, 4]) planned once, captured as a HIP graph and replayed;
  every product of every replay equal to libgmp's (the reference's
  differential tests, test/test_fpntt.jl:288-331, at the bench's shapes).
- hgcd: all 1024 cofactor steps (n log-uniform in [600, 32768], n/2-limb
  cofactors) in adjust mode, both signed outputs exact (the M^-1 application
  of hgcd_matrix_adjust!, src/gcd.jl:398-437).
- serve: Poisson arrivals at 0.5x and 0.8x of the measured capacity, sizes
  log-uniform in [152, 262144], 30% squares, 20% unbalanced (ratio <= 8),
  through the library's queue (nbh_server_*): every result equal to libgmp,
  and a sample equal to the same product computed alone.
"""
import sys

import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip import serve  # noqa: E402


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    assert gmp_ref.available(), "the full-size checks compare against libgmp"
    engine.init()


def _bits(a):
    return engine.magnitude_bits(a, 0, a.size)


def test_batch_config_full_size_graph_replays():
    jobs = wl.batch_jobs(4096)
    ops = wl.batch_operands(jobs)
    say(f"batch: {len(jobs)} jobs, largest {max(j.m for j in jobs)} limbs")
    want = gmp_ref.products(ops)
    tens = []
    for (a, b), w in zip(ops, want):
        ba = _bits(a)
        tens.append((D.empty_limbs(w.size), D.to_dev(a), None if b is None else D.to_dev(b), ba,
                     ba if b is None else _bits(b)))
    B = D.Batch(tens)
    try:
        B.capture()
        for rep in range(2):
            for r, *_ in tens:
                r.fill_(-1)
            B.replay()
            torch.cuda.synchronize()
            bad = [i for i, ((r, *_), w) in enumerate(zip(tens, want)) if not np.array_equal(D.to_host(r), w)]
            say(f"batch replay {rep}: {len(jobs) - len(bad)}/{len(jobs)} equal to libgmp")
            assert not bad, bad[:10]
    finally:
        B.close()


def test_hgcd_config_full_size_adjust():
    sizes = wl.hgcd_sizes(1024)
    host = [wl.hgcd_instance(k, n) for k, n in enumerate(sizes)]
    say(f"hgcd: {len(host)} instances, n up to {max(sizes)}")
    dev = [{k: D.to_dev(v) for k, v in q.items()} for q in host]
    H = D.HgcdStep(dev, "adjust")
    try:
        H.run()
        torch.cuda.synchronize()
        pairs = []
        for q in host:
            pairs += [(q["m11"], q["u"]), (q["m01"], q["v"]), (q["m00"], q["v"]), (q["m10"], q["u"])]
        prods = [wl.limbs_to_int(p) for p in gmp_ref.products(pairs)]
        bad = []
        for i, (r0, r1, sg, _) in enumerate(H.out):
            p = prods[4 * i:4 * i + 4]
            w0, w1 = p[0] - p[1], p[2] - p[3]
            s = sg.cpu().numpy()
            ok = (int(s[0]) == (w0 > 0) - (w0 < 0) and int(s[1]) == (w1 > 0) - (w1 < 0) and
                  int(s[0]) * wl.limbs_to_int(D.to_host(r0)) == w0 and
                  int(s[1]) * wl.limbs_to_int(D.to_host(r1)) == w1)
            if not ok:
                bad.append(i)
        say(f"hgcd: {len(host) - len(bad)}/{len(host)} signed cofactor steps exact")
        assert not bad, bad[:10]
    finally:
        H.close()


@pytest.fixture(scope="module")
def serve_pool():
    pool = 400
    jobs = wl.serve_jobs(pool)
    ops = wl.batch_operands(jobs, workload="serve")
    dev = [(D.to_dev(a), None if b is None else D.to_dev(b)) for a, b in ops]
    bits = [(_bits(a), _bits(a) if b is None else _bits(b)) for a, b in ops]
    want = gmp_ref.products(ops)
    limbs = np.array([w.size for w in want], dtype=np.float64)
    # capacity: the whole pool as one planned batch
    tens = [(D.empty_limbs(int(limbs[i])), dev[i][0], dev[i][1], *bits[i]) for i in range(pool)]
    B = D.Batch(tens)
    try:
        B.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            B.run()
        e1.record()
        torch.cuda.synchronize()
    finally:
        B.close()
    cap = limbs.sum() / (e0.elapsed_time(e1) / 3 * 1e-3)
    say(f"serve pool: {pool} pairs, largest {max(j.m for j in jobs)} limbs, capacity {cap:.4g} limb-products/s")
    return ops, dev, bits, want, limbs, cap


@pytest.mark.parametrize("load", [0.5, 0.8])
def test_serve_config_poisson(serve_pool, load):
    ops, dev, bits, want, limbs, cap = serve_pool
    count = 3000
    rng = np.random.default_rng(int(load * 100))
    pick = rng.integers(0, len(ops), count)
    arrivals = serve.poisson_arrivals(load * cap / limbs[pick].mean(), count, seed=int(load * 1000))
    res = limbs[pick].astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(res)[:-1]])
    out = torch.full((int(res.sum()),), -1, dtype=torch.int64, device="cuda")
    reqs = []
    for p, o in zip(pick.tolist(), offs.tolist()):
        r = serve.Request(dev[p][0], dev[p][1], *bits[p])
        r.out = out[o:o + r.limbs]
        reqs.append(r)
    srv = serve.NativeServer(budget_bytes=8 << 30, nstreams=4, max_batch=256)
    try:
        t0, t_end = serve.drive(srv, reqs, arrivals)
        assert not srv.rejected
        assert sorted(r.rid for r in srv.done) == list(range(count))
        batches = srv.batches
    finally:
        srv.close()
    lat = serve.latency_stats(srv.done)
    say(f"serve {load}: {count} requests in {batches} micro-batches, p50 {lat['p50_ms']:.2f} ms, "
        f"p99 {lat['p99_ms']:.2f} ms, {limbs[pick].sum() / (t_end - t0):.4g} limb-products/s")
    assert batches > 1
    host = out.cpu().numpy().view(np.uint64)
    bad = [i for i, (p, o) in enumerate(zip(pick.tolist(), offs.tolist()))
           if not np.array_equal(host[o:o + want[p].size], want[p])]
    assert not bad, bad[:10]
    # a sample computed alone equals the same request served in the mix
    for i in range(0, count, count // 16):
        p = int(pick[i])
        alone = D.empty_limbs(int(limbs[p]))
        if dev[p][1] is None:
            D.sqr_dev(alone, dev[p][0], bits[p][0])
        else:
            D.mul_dev(alone, dev[p][0], dev[p][1], *bits[p])
        torch.cuda.synchronize()
        assert np.array_equal(D.to_host(alone), host[offs[i]:offs[i] + int(limbs[p])])
