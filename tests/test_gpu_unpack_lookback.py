"""Single products whose unpack resolves the carries between its chunks
itself (at most 256 chunks of 1020 result limbs: published chunk summaries
and a bounded look-back, no k_carry_fix launch) and the first size past that
(the separate carry launch), on both sides of the boundary: uniform,
all-ones and 2^k +- 1 operands (the reference's operand styles, whose carry
chains run through whole chunks), host and device buffers, products and a
square, each checked against libgmp; the per-kernel profile shows which
path ran."""
import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

CHUNK = 1020  # result limbs per unpack chunk (255 threads x 4 limbs)
# (m, n, result chunks); n = 0: the square of a
SHAPES = [(1500, 1500, 3), (130050, 130050, 255), (130560, 130560, 256), (130560, 0, 256), (130560, 130561, 257)]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()
    engine.init()


@pytest.mark.parametrize("m,n,chunks", SHAPES)
def test_products_at_the_lookback_boundary(m, n, chunks):
    rng = np.random.default_rng(m + 7 * n)
    assert -(-(m + (n or m)) // CHUNK) == chunks
    D.profile_enable(True)
    try:
        for style in (1, 2, 3):
            a = rand_big_style(rng, m, style)
            b = None if n == 0 else rand_big_style(rng, n, style)
            want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
            D.profile_read()
            r = np.zeros(want.size, dtype=np.uint64)
            if b is None:
                engine.sqr_fpntt2(r, 0, a, 0, m)
            else:
                engine.mul_fpntt2(r, 0, a, 0, m, b, 0, n)
            prof = D.profile_read()
            assert np.array_equal(r, want), (m, n, style, "host")
            assert prof["unpack"][1] == 1, prof
            assert prof["carry_fix"][1] == (1 if chunks > 256 else 0), prof
            rd = D.empty_limbs(want.size)
            if b is None:
                D.sqr_dev(rd, D.to_dev(a), 64 * m)
            else:
                D.mul_dev(rd, D.to_dev(a), D.to_dev(b), 64 * m, 64 * n)
            torch.cuda.synchronize()
            assert np.array_equal(D.to_host(rd), want), (m, n, style, "device")
    finally:
        D.profile_enable(False)
