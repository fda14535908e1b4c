"""Single products run whole in one workgroup (k_small): pack, both primes'
forward transforms, pointwise products, inverse transforms, Garner CRT and
the carried limb sums in LDS, one launch per call.  It takes every single
product whose two primes' arrays and twiddle tables fit one block's LDS
(transform lengths up to 3072 points).

Checked bit for bit against the oracle (the reference's pipeline restated,
src/fpntt.jl:982-1037 with the unpack of :806-925) and libgmp at every
admissible length it takes, over the reference's operand styles (all-ones
operands give the longest carry chains), for products, squares, unbalanced
shapes and low short products (src/mul.jl:305-378), on host and device
buffers; the per-kernel profile shows which kernel ran."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gmp_ref
from tests.helpers import rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()
    engine.init()


def _N(a, b):
    ba = engine.magnitude_bits(a, 0, a.size)
    bb = ba if b is None else engine.magnitude_bits(b, 0, b.size)
    return engine.fp_ntt_params(ba, bb)[1]


def _shapes():
    """One balanced product per transform length from 512 to 4096 points (the
    small kernel's range and the first length past it), plus squares and
    unbalanced shapes at some of them."""
    seen = {}
    for m in range(150, 1500, 7):
        a = np.full(m, 2**64 - 1, dtype=np.uint64)
        N = _N(a, a)
        seen.setdefault(N, m)
    out = []
    for N, m in sorted(seen.items()):
        out.append((m, m))
        out.append((m, 0))
        out.append((2 * m - 40, m // 3 + 1))
    return out


SHAPES = _shapes()


def _run_host(a, b):
    m = a.size
    if b is None:
        r = np.zeros(2 * m, dtype=np.uint64)
        engine.sqr_fpntt2(r, 0, a, 0, m)
    else:
        r = np.zeros(m + b.size, dtype=np.uint64)
        engine.mul_fpntt2(r, 0, a, 0, m, b, 0, b.size)
    return r


def test_routing_by_profile():
    """Lengths whose arrays fit one block run as one small launch and nothing
    else; the next length up takes the tiled pipeline."""
    rng = np.random.default_rng(5)
    D.profile_enable(True)
    try:
        D.profile_read()
        a = rand_big_style(rng, 700, 1)
        b = rand_big_style(rng, 650, 1)
        assert _N(a, b) <= 3072
        assert np.array_equal(_run_host(a, b), gmp_ref.mul(a, b))
        prof = D.profile_read()
        assert prof["small"][1] == 1, prof
        assert prof["mid"][1] == 0 and prof["unpack"][1] == 0, prof
        a = rand_big_style(rng, 1500, 1)
        b = rand_big_style(rng, 1500, 1)
        assert _N(a, b) > 3072
        assert np.array_equal(_run_host(a, b), gmp_ref.mul(a, b))
        prof = D.profile_read()
        assert prof["small"][1] == 0 and prof["mid"][1] >= 1, prof
    finally:
        D.profile_enable(False)


@pytest.mark.parametrize("style", [1, 2, 3, 4])
def test_host_products_every_length(style):
    rng = np.random.default_rng(300 + style)
    for m, n in SHAPES:
        a = rand_big_style(rng, m, style)
        b = None if n == 0 else rand_big_style(rng, n, style)
        got = _run_host(a, b)
        want = O.sqr(a) if b is None else O.mul(a, b)
        assert np.array_equal(got, want), (m, n, style, _N(a, b))
        assert np.array_equal(want, gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b))


def test_device_products_and_zero_tops():
    rng = np.random.default_rng(11)
    for m, n in SHAPES[::2]:
        a = rand_big_style(rng, m, 1)
        b = None if n == 0 else rand_big_style(rng, n, 1)
        r = torch.full((2 * m if b is None else m + n,), -1, dtype=torch.int64, device="cuda")
        if b is None:
            D.sqr_dev(r, D.to_dev(a))
        else:
            D.mul_dev(r, D.to_dev(a), D.to_dev(b))
        torch.cuda.synchronize()
        want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
        assert np.array_equal(D.to_host(r), want), (m, n)
    # zero top limbs: the magnitude, not the buffer, sizes the transform
    a = rand_big_style(rng, 900, 1)
    a[-300:] = 0
    b = rand_big_style(rng, 500, 1)
    b[-7:] = 0
    r = D.empty_limbs(1400)
    D.mul_dev(r, D.to_dev(a), D.to_dev(b))
    torch.cuda.synchronize()
    assert np.array_equal(D.to_host(r), gmp_ref.mul(a, b))


def test_short_products():
    rng = np.random.default_rng(17)
    for la, lb, k in [(300, 250, 300), (300, 250, 301), (300, 250, 549), (300, 250, 550), (700, 0, 700),
                      (700, 0, 1399), (1000, 900, 1000), (1000, 900, 1901), (400, 399, 799)]:
        a = rand_big_style(rng, la, 2 if k % 2 else 1)
        b = None if lb == 0 else rand_big_style(rng, lb, 2 if k % 2 else 1)
        r = np.full(k, 7, dtype=np.uint64)
        if b is None:
            engine.sqrlo(r, 0, a, 0, la, k)
        else:
            engine.mullo(r, 0, a, 0, la, b, 0, lb, k)
        full = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
        want = np.zeros(k, dtype=np.uint64)
        want[:min(k, full.size)] = full[:k]
        assert np.array_equal(r, want), (la, lb, k)


def test_repeated_calls_alternate_with_tiled_products():
    """Small and tiled single products in turn share the library's ordering
    and buffers: every result stays exact."""
    rng = np.random.default_rng(23)
    cases = [(400, 380), (2500, 2500), (1000, 0), (6000, 3000), (160, 160), (1200, 1100)]
    for _ in range(3):
        for m, n in cases:
            a = rand_big_style(rng, m, 1)
            b = None if n == 0 else rand_big_style(rng, n, 1)
            assert np.array_equal(_run_host(a, b), gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)), (m, n)
