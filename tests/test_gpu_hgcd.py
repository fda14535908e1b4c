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
