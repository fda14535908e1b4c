"""Host-buffer batches (nbh_mul_batch): many independent mul!/sqr!/mullo!/
sqrlo! calls (src/mul.jl:127-164, 305-389) in one call, every result
bit-exact against libgmp, with sentinels around each result to show that
exactly its limbs are written.  Covers the schoolbook and NTT ranges, squares,
unbalanced pairs, short products, operands with zero top limbs, zero operands,
and batches larger than the staging group (several groups, and a job larger
than a group, copied directly)."""
import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import engine  # noqa: E402
from nbhip._lib import NbhError  # noqa: E402

SENT = 0xDEADBEEFDEADBEEF
GAP = 3  # sentinel limbs between results


@pytest.fixture(scope="module", autouse=True)
def _init():
    if not gmp_ref.available():
        pytest.skip("libgmp not found")
    engine.init()


def _want(a, b, k):
    """The product, or its low k limbs (zero-padded when k > m + n)."""
    full = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
    if not k:
        return full
    out = np.zeros(k, dtype=np.uint64)
    out[:min(k, full.size)] = full[:k]
    return out


def _run(specs, rng):
    """specs: (m, n or None, k, style); returns (jobs, expected, r, offsets)."""
    ops, sizes = [], []
    for m, n, k, style in specs:
        a = rand_big_style(rng, m, style)
        b = None if n is None else rand_big_style(rng, n, style)
        if style == 5:            # not top-normalised (magnitude_bits stays an upper bound)
            a[-1] = 0
        if style == 6:            # a zero operand
            a[:] = 0
        ops.append((a, b))
        mm = min(m, k) if k else m
        nn = mm if n is None else (min(n, k) if k else n)
        sizes.append(k if k else mm + nn)
    offs = np.cumsum([GAP] + [s + GAP for s in sizes])[:-1]
    r = np.full(int(offs[-1] + sizes[-1] + GAP), SENT, dtype=np.uint64)
    jobs = []
    for (a, b), (m, n, k, _), o in zip(ops, specs, offs):
        jobs.append((r, int(o), a, 0, m, b, 0, 0 if n is None else n, k))
    engine.mul_batch(jobs)
    gaps = np.ones(r.size, dtype=bool)
    for (a, b), (m, n, k, _), o, s in zip(ops, specs, offs, sizes):
        aa = a[:k] if k else a
        bb = None if b is None else (b[:k] if k else b)
        want = _want(aa, bb, k)
        got = r[o:o + s]
        assert np.array_equal(got, want), (m, n, k)
        gaps[o:o + s] = False
    assert (r[gaps] == SENT).all()


def test_mixed_sizes_and_kinds():
    rng = np.random.default_rng(0xB47C)
    specs = []
    for _ in range(160):
        m = int(np.exp(rng.uniform(0, np.log(20000))))
        kind = int(rng.integers(0, 4))
        style = int(rng.integers(1, 7))
        if kind == 0:
            specs.append((m, None, 0, style))
        elif kind == 1:
            specs.append((m, max(1, int(m / rng.uniform(1, 6))), 0, style))
        elif kind == 2:
            specs.append((m, int(rng.integers(1, m + 1)), int(rng.integers(1, 2 * m + 1)), style))
        else:
            specs.append((m, None, int(rng.integers(1, 2 * m + 1)), style))
    # the routing edges of mul!/sqr! (src/mul.jl:24-26)
    specs += [(152, 152, 0, 1), (151, 151, 0, 2), (240, 64, 0, 1), (241, 63, 0, 2),
              (160, None, 0, 2), (159, None, 0, 1), (1, 1, 0, 2), (1, None, 0, 6)]
    _run(specs, rng)


def test_matches_single_products():
    rng = np.random.default_rng(7)
    a = rand_big_style(rng, 4096, 1)
    b = rand_big_style(rng, 3000, 1)
    r1 = np.zeros(7096, dtype=np.uint64)
    engine.mul_fpntt2(r1, 0, a, 0, 4096, b, 0, 3000)
    r2 = np.zeros(7096 + 8192, dtype=np.uint64)
    engine.mul_batch([(r2, 0, a, 0, 4096, b, 0, 3000, 0), (r2, 7096, a, 0, 4096, None, 0, 0, 0)])
    assert np.array_equal(r1, r2[:7096])
    r3 = np.zeros(8192, dtype=np.uint64)
    engine.sqr_fpntt2(r3, 0, a, 0, 4096)
    assert np.array_equal(r3, r2[7096:])


def test_several_groups_and_an_oversized_job():
    """Eighteen 2^19 x 2^19 products (16 MiB of operands and result each)
    need three 128 MiB staging groups, so a staging slot is reused; a
    12M-limb square (288 MiB) runs on its own, copied directly."""
    rng = np.random.default_rng(0x6B)
    specs = [(1 << 19, 1 << 19, 0, 1)] * 18 + [(12 << 20, None, 0, 2), (300, 200, 0, 1)]
    _run(specs, rng)


def test_errors():
    a = np.ones(200, dtype=np.uint64)
    r = np.zeros(400, dtype=np.uint64)
    with pytest.raises(ValueError):
        engine.mul_batch([(r, 0, a, 0, 200, a, 0, 200, 0), (r, 399, a, 0, 1, None, 0, 0, 1)])
    with pytest.raises(ValueError):
        engine.mul_batch([(r, 0, r, 100, 200, a, 0, 200, 0)])
    engine.mul_batch([])
    with pytest.raises(NbhError):
        from nbhip._lib import lib
        from nbhip._lib import check
        check(lib().nbh_mul_batch(None, 0))
