"""Single products captured into a HIP graph (the callers the allocation-free
entry point is for): nbh_mul_dev_noalloc recorded once on a torch stream and
replayed with new operands, with the program's own eager products and
synchronisations between the replays; every replay is checked against the
eager result and libgmp.  Covers the one-workgroup kernel, a tiled product
whose unpack resolves its own carries (chunk flags cleared by the mid launch
in every replay) and a longer tiled product."""
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


@pytest.mark.parametrize("m,n", [(300, 200), (4096, 4096), (9000, 7000)])
def test_captured_single_product_replays(m, n):
    engine.init()
    a0, b0 = wl.operand("test", 1, 0, m), wl.operand("test", 1, 1, n)
    da, db = D.to_dev(a0), D.to_dev(b0)
    r = D.empty_limbs(m + n)
    D.mul_dev(r, da, db, 64 * m, 64 * n)          # caches the plans the capture needs
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            rc = _noalloc(r, da, db, 64 * m, 64 * n)
    assert rc == 0, _lib.lib().nbh_last_error_string()
    for rep in range(4):
        a, b = wl.operand("test", 10 + rep, 0, m), wl.operand("test", 10 + rep, 1, n)
        da.copy_(D.to_dev(a))
        db.copy_(D.to_dev(b))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        eager = D.empty_limbs(m + n)
        D.mul_dev(eager, da, db, 64 * m, 64 * n)
        torch.cuda.synchronize()
        got = D.to_host(r)
        assert np.array_equal(got, D.to_host(eager)), (m, n, rep)
        assert np.array_equal(got, gmp_ref.mul(a, b)), (m, n, rep)
