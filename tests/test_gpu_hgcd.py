"""Batched 2x2 cofactor step (the arithmetic of hgcd_mul2! and
hgcd_matrix_adjust!, src/gcd.jl:360-437) against Python big ints."""
import numpy as np
import pytest

from tests.helpers import ints, wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402


def _instance(seed, n, mn, pad=0, equal=False):
    rng = np.random.default_rng(seed)

    def limbs(k, top_zero=0):
        a = rng.integers(0, 2**64, k, dtype=np.uint64)
        if top_zero:
            a[-top_zero:] = 0
        return a
    q = {"m00": limbs(mn, pad), "m01": limbs(mn), "m10": limbs(mn), "m11": limbs(mn, pad),
         "u": limbs(n), "v": limbs(n)}
    if equal:                       # m11*u == m01*v and m00*v == m10*u
        q["m01"] = q["m11"].copy()
        q["v"] = q["u"].copy()
        q["m10"] = q["m00"].copy()
    return q


CASES = [(1, 600, 300), (2, 1000, 500), (3, 2500, 1250), (4, 700, 350, 40), (5, 333, 200),
         (6, 900, 450, 0, True), (7, 4097, 2049), (8, 160, 80), (9, 12000, 6000, 100)]


@pytest.fixture(scope="module")
def instances():
    engine.init()
    host = [_instance(*c[:1], *c[1:3], *(c[3:] if len(c) > 3 else ())) for c in CASES]
    dev = [{k: D.to_dev(v) for k, v in q.items()} for q in host]
    return host, dev


def test_forward(instances):
    host, dev = instances
    H = D.HgcdStep(dev, "forward")
    H.run()
    torch.cuda.synchronize()
    for (r0, r1, _, _), q in zip(H.out, host):
        M = {k: ints(v) for k, v in q.items()}
        assert ints(D.to_host(r0)) == M["m00"] * M["u"] + M["m01"] * M["v"]
        assert ints(D.to_host(r1)) == M["m10"] * M["u"] + M["m11"] * M["v"]
    H.close()


def test_adjust_signed(instances):
    host, dev = instances
    H = D.HgcdStep(dev, "adjust")
    for rep in range(2):            # a planned step can be re-run
        H.run()
        torch.cuda.synchronize()
        for (r0, r1, sg, _), q in zip(H.out, host):
            M = {k: ints(v) for k, v in q.items()}
            s = sg.cpu().numpy()
            w0 = M["m11"] * M["u"] - M["m01"] * M["v"]
            w1 = M["m00"] * M["v"] - M["m10"] * M["u"]
            assert int(s[0]) * ints(D.to_host(r0)) == w0
            assert int(s[1]) * ints(D.to_host(r1)) == w1
            assert int(s[0]) == (w0 > 0) - (w0 < 0) and int(s[1]) == (w1 > 0) - (w1 < 0)
    H.close()


def test_matrix_mul():
    """hgcd_matrix_mul! (src/gcd.jl:372-391): M <- M * M1 for a batch of
    matrix pairs, as two forward cofactor steps per pair."""
    rng = np.random.default_rng(44)

    def mat(k, pad=0):
        out = {}
        for name in ("m00", "m01", "m10", "m11"):
            a = rng.integers(0, 2**64, k, dtype=np.uint64)
            if pad:
                a[-pad:] = 0
            out[name] = a
        return out
    shapes = [(300, 200), (1000, 1000), (160, 2500), (4000, 90, 3)]
    host = [(mat(n), mat(n1, *rest)) for n, n1, *rest in shapes]
    pairs = [({k: D.to_dev(v) for k, v in M.items()}, {k: D.to_dev(v) for k, v in M1.items()}) for M, M1 in host]
    H = D.HgcdMatrixMul(pairs)
    H.run()
    torch.cuda.synchronize()
    for i, (M, M1) in enumerate(host):
        A = {k: ints(v) for k, v in M.items()}
        B = {k: ints(v) for k, v in M1.items()}
        R = H.result(i)
        assert ints(D.to_host(R["m00"])) == A["m00"] * B["m00"] + A["m01"] * B["m10"]
        assert ints(D.to_host(R["m01"])) == A["m00"] * B["m01"] + A["m01"] * B["m11"]
        assert ints(D.to_host(R["m10"])) == A["m10"] * B["m00"] + A["m11"] * B["m10"]
        assert ints(D.to_host(R["m11"])) == A["m10"] * B["m01"] + A["m11"] * B["m11"]
        n, n1 = M["m00"].size, M1["m00"].size
        assert R["m00"].numel() == n + n1 + 1
    H.close()


# (seed, p, mn, nn, force the first result negative); with nn = 1 the top
# term is small and the results come out negative on their own
FULL_CASES = [(21, 600, 300, 320, False), (22, 1000, 500, 480, False), (23, 333, 200, 1, False),
              (24, 2500, 1250, 1300, False), (25, 700, 350, 351, True), (26, 160, 80, 90, False),
              (27, 12000, 6000, 6100, False), (28, 5, 3, 2, False)]


@pytest.mark.parametrize("seed,p,mn,nn,neg", FULL_CASES)
def test_adjust_full(seed, p, mn, nn, neg):
    """hgcd_matrix_adjust! (src/gcd.jl:398-437): u' = ra*2^(64p) + m11*ulow
    - m01*vlow, v' = rb*2^(64p) + m00*vlow - m10*ulow, zero-padded to
    L = p + max(mn+1, nn) + 1 limbs; a result the reference would assert on
    (negative) comes back with sign -1 and its magnitude."""
    engine.init()
    rng = np.random.default_rng(seed)
    q = {k: rng.integers(0, 2**64, mn, dtype=np.uint64) for k in ("m00", "m01", "m10", "m11")}
    q["u"] = rng.integers(0, 2**64, p, dtype=np.uint64)
    q["v"] = rng.integers(0, 2**64, p, dtype=np.uint64)
    q["ra"] = rng.integers(0, 2**64, nn, dtype=np.uint64)
    q["rb"] = rng.integers(0, 2**64, nn, dtype=np.uint64)
    q["rb"][-1] |= np.uint64(1 << 63)
    if neg:
        q["ra"][:] = 0                # m01*vlow > m11*ulow makes u' negative
        q["m11"][-1] = np.uint64(1)
        q["m01"][-1] = np.uint64(2**64 - 1)
    M = {k: ints(v) for k, v in q.items()}
    L = p + max(mn + 1, nn) + 1
    e0 = (M["ra"] << (64 * p)) + M["m11"] * M["u"] - M["m01"] * M["v"]
    e1 = (M["rb"] << (64 * p)) + M["m00"] * M["v"] - M["m10"] * M["u"]
    H = D.HgcdStep([{k: D.to_dev(v) for k, v in q.items()}], "adjust_full")
    for rep in range(2):
        for t in H.out[0][:2]:
            t.fill_(-1)
        H.run()
        torch.cuda.synchronize()
        r0, r1, sg, _ = H.out[0]
        assert r0.numel() == L and r1.numel() == L
        s = sg.cpu().numpy()
        for r, e, sign in ((r0, e0, int(s[0])), (r1, e1, int(s[1]))):
            assert sign == (1 if e > 0 else 0 if e == 0 else -1)
            assert ints(D.to_host(r)) == abs(e)
    if neg:
        assert e0 < 0               # the case meant to exercise the assertion's sign
    H.close()
