"""nbh_mul_dev_noalloc (the product without any allocation, in the library's
static device workspace) against nbh_mul_dev, the entry point it is a variant
of, and libgmp: bit-exact on two calls in a row, the second smaller, and a
product beyond the workspace refused before any launch."""
import ctypes

import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import _lib  # noqa: E402
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402


def _noalloc(r, a, b, bits_a, bits_b):
    return _lib.lib().nbh_mul_dev_noalloc(D._p(r), D._p(a), a.numel(), D._p(b), b.numel(), bits_a, bits_b,
                                          D._stream())


def test_noalloc_equals_mul_dev_twice_second_smaller():
    engine.init()
    shapes = [(9000, 7000), (2500, 700)]         # the second call smaller (and a shorter transform)
    cases = []
    for k, (m, n) in enumerate(shapes):
        a, b = wl.operand("test", 800 + k, 0, m), wl.operand("test", 800 + k, 1, n)
        da, db = D.to_dev(a), D.to_dev(b)
        ba, bb = engine.magnitude_bits(a, 0, m), engine.magnitude_bits(b, 0, n)
        ref = D.empty_limbs(m + n)
        D.mul_dev(ref, da, db, ba, bb)           # the existing entry point (and it caches the plans)
        cases.append((a, b, da, db, ba, bb, ref))
    torch.cuda.synchronize()
    outs = []
    for a, b, da, db, ba, bb, ref in cases:
        r = torch.full((a.size + b.size,), -1, dtype=torch.int64, device="cuda")
        _lib.check(_noalloc(r, da, db, ba, bb))
        outs.append(r)
    torch.cuda.synchronize()
    for (a, b, _, _, _, _, ref), r in zip(cases, outs):
        got = D.to_host(r)
        assert np.array_equal(got, D.to_host(ref))
        assert np.array_equal(got, gmp_ref.mul(a, b))


def test_noalloc_refuses_products_beyond_its_workspace():
    engine.init()
    m = n = 1 << 20                               # 4 spectra of 7 * 2^19 doubles: 112 MiB
    a = torch.ones(m, dtype=torch.int64, device="cuda")
    b = torch.ones(n, dtype=torch.int64, device="cuda")
    r = torch.full((m + n,), -1, dtype=torch.int64, device="cuda")
    rc = _noalloc(r, a, b, 64 * m, 64 * n)
    assert rc == _lib.NBH_ERR_ARG
    assert b"workspace" in _lib.lib().nbh_last_error_string()
    torch.cuda.synchronize()
    assert bool((r == -1).all())                  # refused before any launch
