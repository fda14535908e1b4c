"""GPU parity: the HIP engine through the C ABI against the oracle, libgmp
and the committed fixtures.  Bit-exact everywhere: spectra (the transform
network is the reference's butterfly for butterfly), twiddles, constants and
product limbs.  Shapes follow test/test_fpntt.jl and test/test_algorithms.jl:
109-128."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tests import gmp_ref
from tests.helpers import ints, rand_big_style, wl

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

P = (65205 * 2**34 + 1, 4095 * 2**38 + 1)
FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "products.json")))
DEFAULT_TUNING = None


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    global DEFAULT_TUNING
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    engine.init()
    DEFAULT_TUNING = D.get_tuning()
    say(f"device {D.device_info()} tuning {DEFAULT_TUNING}")
    yield
    D.set_tuning(*DEFAULT_TUNING)


def sha(r):
    return hashlib.sha256(np.ascontiguousarray(r, dtype="<u8").tobytes()).hexdigest()


LENGTHS = [4, 8, 16, 32, 64, 512, 1024, 12, 20, 48, 96, 160, 240, 1536, 960, 3840, 28, 56, 84, 140, 420, 1680,
           36, 144, 720, 1260, 4032, 12288, 16384, 7 * 2**12, 3 * 2**13, 45 * 2**10, 315 * 2**8, 2**16,
           105 * 2**10, 63 * 2**12, 21 * 2**13, 35 * 2**12, 15 * 2**14, 9 * 2**15, 7 * 2**16, 2**18, 5 * 2**17]


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", [12, 28, 36, 1260, 4032, 12288, 7 * 2**16, 315 * 2**8])
def test_plan_constants_match_oracle(k, N):
    say(f"plan constants N={N} prime={k}")
    ref = O.Plan(N, k)
    got = D.plan_stages(N, k)
    radix2 = (ref.N2.bit_length() - 1) % 2 == 1
    assert len(got) == len(ref.stages) + (1 if radix2 else 0)
    for i, (m, span, Q, rot) in enumerate(ref.stages):
        gm, gspan, gQ, grot = got[i]
        assert (gm, gQ) == (m, Q)
        nrot = {3: 2, 4: 1, 5: 4, 7: 8}[m]
        assert np.array_equal(grot[:nrot], rot[:nrot])
        for t in range(1, m):
            assert np.array_equal(D.plan_twiddle_row(N, k, i, t, Q), ref.twiddle_row(i, t - 1)), (i, t)
    assert D.plan_ninv(N, k) == (ref.ninv, ref.ninvp)


def _transform_case(N, k, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, P[k], N).astype(np.float64)
    ref = O.Plan(N, k)
    want_f = ref.fwd(x)
    t = torch.from_numpy(x).cuda()
    D.fp_ntt_fwd(t, k)
    got_f = t.cpu().numpy()
    assert np.array_equal(got_f, want_f), f"forward N={N} prime={k}: {np.sum(got_f != want_f)} mismatches"
    want_r = ref.rev(want_f)
    D.fp_ntt_rev(t, k)
    got_r = t.cpu().numpy()
    assert np.array_equal(got_r, want_r), f"reverse N={N} prime={k}"
    # round trip identity N * x[(N-t) mod N] (test/test_fpntt.jl:84-99)
    idx = (N - np.arange(N)) % N
    assert all((int(v) - N * int(x[i])) % P[k] == 0 for v, i in zip(got_r[::max(1, N // 64)], idx[::max(1, N // 64)]))


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", LENGTHS)
def test_transform_bit_exact(k, N):
    say(f"transform N={N} prime={k}")
    _transform_case(N, k, N * 3 + k)


@pytest.mark.parametrize("tuning", [(256, 64, 1, 4), (64, 16, 2, 2), (1024, 256, 4, 16), (16384, 8192, 8, 64),
                                    (4096, 16384, 8, 64)])
@pytest.mark.parametrize("N", [12288, 7 * 2**12, 315 * 2**8, 105 * 2**10, 2**16, 9 * 2**12, 5 * 2**13])
def test_transform_other_schedules(tuning, N):
    """Several strided groups (odd radices inside them, narrow and wide column
    tiles) must not change a single bit."""
    say(f"schedule {tuning} N={N}")
    try:
        D.set_tuning(*tuning)
    except Exception as e:  # a tuning beyond this device's LDS
        pytest.skip(str(e))
    D.set_tuning_table([])          # the autotuned per-length table would override it
    try:
        assert D.get_tuning_for(N) == tuning
        say(f"  groups {D.plan_groups(N)}")
        for k in (0, 1):
            _transform_case(N, k, N + 17 * k)
    finally:
        D.set_tuning(*DEFAULT_TUNING)
        D.set_tuning_table(D.tuning_file_entries())


@pytest.mark.parametrize("n,b,nch,N", [(1, 49, 2, 4), (5, 48, 7, 12), (152, 46, 212, 512), (4096, 43, 6097, 12288),
                                       (300, 45, 427, 768), (3, 33, 6, 8)])
def test_pack_bit_exact(n, b, nch, N):
    limbs = wl.operand("test", 100 + n, 0, n)
    want = O.pack(limbs, b, nch, N)
    x = torch.full((N,), 7.0, dtype=torch.float64, device="cuda")
    D.fp_ntt_pack(x, D.to_dev(limbs), b, nch)
    assert np.array_equal(x.cpu().numpy(), want)


@pytest.mark.parametrize("k", [0, 1])
def test_pointwise_bit_exact(k):
    rng = np.random.default_rng(5 + k)
    xa = (rng.integers(-P[k], P[k], 10007)).astype(np.float64)
    xb = (rng.integers(-P[k], P[k], 10007)).astype(np.float64)
    ta, tb = torch.from_numpy(xa).cuda(), torch.from_numpy(xb).cuda()
    D.fp_ntt_pointwise(ta, tb, k)
    assert np.array_equal(ta.cpu().numpy(), O.pointwise(xa, xb, k))


@pytest.mark.parametrize("b", [33, 40, 44, 48, 49])
@pytest.mark.parametrize("nconv", [1, 7, 64, 200, 5000])
def test_unpack_vs_crt(b, nconv):
    """Garner + carries against big-int CRT (test/test_fpntt.jl:257-286); the
    layout is what fp_ntt_rev! hands over: coefficient i at (N-i) mod N, times N."""
    say(f"unpack b={b} nconv={nconv}")
    rng = np.random.default_rng(0x6A3 + b * 1000 + nconv)
    N = O.ntt_len(nconv)
    cs = [int.from_bytes(rng.bytes(16), "little") % (P[0] * P[1]) for _ in range(nconv)]
    x1 = np.zeros(N)
    x2 = np.zeros(N)
    for i, c in enumerate(cs):
        idx = (N - i) % N
        x1[idx] = float((c * N) % P[0]) + int(rng.integers(-2, 3)) * float(P[0])
        x2[idx] = float((c * N) % P[1]) + int(rng.integers(-2, 3)) * float(P[1])
    rn = -(-(99 + b * (nconv - 1)) // 64) + 1
    p1 = O.Plan(N, 0)
    p2 = O.Plan(N, 1)
    want = O.unpack2(rn, x1, x2, nconv, b, p1.ninv, p1.ninvp, p2.ninv, p2.ninvp)
    r = D.empty_limbs(rn)
    D.fp_ntt_unpack2(r, torch.from_numpy(x1).cuda(), torch.from_numpy(x2).cuda(), nconv, b)
    got = D.to_host(r)
    assert np.array_equal(got, want)
    assert ints(got) == sum(c << (b * i) for i, c in enumerate(cs)) % (1 << (64 * rn))


def test_unpack_long_carry_chains():
    """All-ones products make carries ripple across whole chunks and across
    chunk boundaries (the look-back path)."""
    for n in (600, 2100, 4097, 9000):
        a = np.full(n, 2**64 - 1, dtype=np.uint64)
        r = np.empty(2 * n, dtype=np.uint64)
        engine.sqr_fpntt2(r, 0, a, 0, n)
        assert ints(r) == ints(a) ** 2, n
        b = np.full(n - 7, 2**64 - 1, dtype=np.uint64)
        r2 = np.empty(2 * n - 7, dtype=np.uint64)
        engine.mul_fpntt2(r2, 0, a, 0, n, b, 0, n - 7)
        assert ints(r2) == ints(a) * ints(b), n


@pytest.mark.parametrize("na,nb", [(1, 1), (5, 3), (16, 16), (100, 100), (255, 257), (500, 300), (601, 373),
                                   (1024, 1024), (2500, 2500), (5000, 4900), (151, 151), (152, 152), (153, 153),
                                   (456, 76), (381, 153), (64, 240), (3000, 64), (8191, 8193), (20000, 13)])
def test_mul_differential(na, nb):
    say(f"mul {na}x{nb}")
    rng = np.random.default_rng(0x91C4 + na * 7919 + nb)
    for trial in range(3):
        a = rand_big_style(rng, na)
        b = rand_big_style(rng, nb)
        r = np.empty(na + nb, dtype=np.uint64)
        engine.mul_fpntt2(r, 0, a, 0, na, b, 0, nb)
        want = O.mul(a, b)
        assert np.array_equal(r, want), (na, nb, trial)
        assert ints(r) == ints(a) * ints(b)


@pytest.mark.parametrize("n", [1, 7, 100, 159, 160, 161, 321, 1024, 3000, 12345])
def test_sqr_differential(n):
    rng = np.random.default_rng(0x3F17 + n)
    for trial in range(3):
        a = rand_big_style(rng, n)
        r = np.empty(2 * n, dtype=np.uint64)
        engine.sqr_fpntt2(r, 0, a, 0, n)
        assert np.array_equal(r, O.sqr(a)), (n, trial)


def test_all_ones_reach_b49_and_zero():
    for na in range(1, 9):
        for nb in range(1, na + 1):
            a = np.full(na, 2**64 - 1, dtype=np.uint64)
            b = np.full(nb, 2**64 - 1, dtype=np.uint64)
            r = np.empty(na + nb, dtype=np.uint64)
            engine.mul_fpntt2(r, 0, a, 0, na, b, 0, nb)
            assert ints(r) == ints(a) * ints(b)
    z = np.zeros(1, dtype=np.uint64)
    a = wl.operand("test", 3, 0, 300)
    r = np.full(301, 5, dtype=np.uint64)
    engine.mul_fpntt2(r, 0, a, 0, 300, z, 0, 1)
    assert not r.any()


def test_offsets_and_dispatch_mirror():
    """mul!/sqr!/mul_bal! at the crossover (test/test_algorithms.jl:109-128),
    with nonzero buffer offsets."""
    T = engine.MUL_FPNTT_THRESHOLD
    rng = np.random.default_rng(41)
    for m, n in [(T, T), (T + 1, T + 1), (3 * T, T // 2), (2 * T + T // 2, T + 1)]:
        a = np.concatenate([rng.integers(0, 2**64, 3, dtype=np.uint64), rng.integers(0, 2**64, m, dtype=np.uint64)])
        b = np.concatenate([rng.integers(0, 2**64, 5, dtype=np.uint64), rng.integers(0, 2**64, n, dtype=np.uint64)])
        r = np.zeros(m + n + 9, dtype=np.uint64)
        engine.mul(r, 9, a, 3, m, b, 5, n)
        assert ints(r[9:]) == ints(a[3:]) * ints(b[5:])
        assert not r[:9].any()
    for n in (engine.SQR_FPNTT_THRESHOLD, 2 * engine.SQR_FPNTT_THRESHOLD + 1):
        a = rng.integers(0, 2**64, n, dtype=np.uint64)
        r = np.empty(2 * n, dtype=np.uint64)
        engine.sqr(r, 0, a, 0, n)
        assert ints(r) == ints(a) ** 2
        r2 = np.empty(2 * n, dtype=np.uint64)
        engine.mul_bal(r2, 0, a, 0, a, 0, n)
        assert np.array_equal(r, r2)


def test_device_path_matches_host_path():
    a = wl.operand("test", 7, 0, 3333)
    b = wl.operand("test", 7, 1, 2222)
    want = O.mul(a, b)
    r = D.empty_limbs(a.size + b.size)
    D.mul_dev(r, D.to_dev(a), D.to_dev(b))
    torch.cuda.synchronize()
    assert np.array_equal(D.to_host(r), want)
    s = D.empty_limbs(2 * a.size)
    D.sqr_dev(s, D.to_dev(a))
    torch.cuda.synchronize()
    assert np.array_equal(D.to_host(s), O.sqr(a))


def test_batch_alone_equals_in_mix():
    """Every product of a mixed batch equals the oracle and the same product
    computed alone (one per launch)."""
    jobs = wl.batch_jobs(count=96, lo=1, hi=20000, workload="test")
    ops = wl.batch_operands(jobs, workload="test")
    tens = []
    for j, (a, b) in zip(jobs, ops):
        r = D.empty_limbs(j.m + (j.m if b is None else j.n))
        tens.append((r, D.to_dev(a), None if b is None else D.to_dev(b)))
    B = D.Batch(tens)
    B.run()
    torch.cuda.synchronize()
    for (r, ta, tb), (a, b) in zip(tens, ops):
        got = D.to_host(r)
        want = O.sqr(a) if b is None else O.mul(a, b)
        assert np.array_equal(got, want)
        alone = D.empty_limbs(r.numel())
        if b is None:
            D.sqr_dev(alone, ta)
        else:
            D.mul_dev(alone, ta, tb)
        torch.cuda.synchronize()
        assert np.array_equal(D.to_host(alone), got)
    # re-running the planned batch gives the same limbs
    for r, _, _ in tens:
        r.fill_(0)
    B.run()
    torch.cuda.synchronize()
    for (r, _, _), (a, b) in zip(tens, ops):
        assert np.array_equal(D.to_host(r), O.sqr(a) if b is None else O.mul(a, b))
    B.close()


@pytest.mark.parametrize("key", [k for k in FIX["products"]])
def test_fixture_products(key):
    """Full-size workloads (1M x 1M, 1M x 256K, 4M square, ...) against the
    libgmp SHA-256 fixtures; the device path, inputs already in HBM."""
    say(f"fixture {key}")
    wname, job, shape = key.split("/")[:3]
    m, n = (int(v) for v in shape.split("x"))
    a = D.to_dev(wl.operand(wname, int(job), 0, m))
    if key.endswith("/sqr"):
        r = D.empty_limbs(2 * m)
        D.sqr_dev(r, a)
    else:
        r = D.empty_limbs(m + n)
        D.mul_dev(r, a, D.to_dev(wl.operand(wname, int(job), 1, n)))
    torch.cuda.synchronize()
    assert sha(D.to_host(r)) == FIX["products"][key]


def test_large_by_residues():
    """A 2^20-limb product checked mod random 63-bit moduli
    (test/test_fpntt.jl:313-324): an O(n) check independent of the engine."""
    n = 1 << 20
    a = wl.operand("test", 30, 0, n)
    b = wl.operand("test", 30, 1, n - 12345)
    r = D.empty_limbs(2 * n - 12345)
    D.mul_dev(r, D.to_dev(a), D.to_dev(b))
    torch.cuda.synchronize()
    got = D.to_host(r)
    rng = np.random.default_rng(11)
    for mod in rng.integers(2**62, 2**63 - 1, 3):
        mod = int(mod)

        def red(x):
            acc = 0
            for v in x[::-1].tolist():
                acc = ((acc << 64) | v) % mod
            return acc
        assert red(got) == red(a) * red(b) % mod


def test_length_error_is_argument_error():
    with pytest.raises(engine.ArgumentError):
        engine.ntt_len((315 << 34) + 1)


def test_single_product_table():
    """Products run alone read the 'single' lines of the schedule table (and
    the batch lines past them); batches keep the batch lines.  Both give the
    same limbs."""
    single = D.tuning_file_entries(single=True)
    if not single:
        pytest.skip("no single-product entries shipped")
    for nmax, knobs in single:
        assert D.get_tuning_for_single(nmax) == knobs
    past = single[-1][0] * 2
    assert D.get_tuning_for_single(past) == D.get_tuning_for(past)
    rng = np.random.default_rng(0x51)
    for n in (2048, 4096, 9000):
        a, b = rng.integers(0, 2**64, n, dtype=np.uint64), rng.integers(0, 2**64, n - 7, dtype=np.uint64)
        r1 = np.zeros(2 * n - 7, dtype=np.uint64)
        engine.mul_fpntt2(r1, 0, a, 0, n, b, 0, n - 7)
        r2 = np.zeros(2 * n - 7, dtype=np.uint64)
        engine.mul_batch([(r2, 0, a, 0, n, b, 0, n - 7, 0)])
        assert np.array_equal(r1, r2)
        assert np.array_equal(r1, O.mul(a, b))
