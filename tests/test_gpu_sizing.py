"""The sizing functions on the GPU box, after nbh_init: ntt_len
(src/fpntt.jl:936-951), fp_ntt_params (:964-977) and magnitude_bits
(src/kernels/addsub.jl:150, read from the top limb on the device) against the
oracle and the reference test's literal values (test/test_fpntt.jl:205-255,
308-316), so the GPU run covers them too."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import gmp_ref
from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()
    engine.init()


KATS_LEN = [(7 << 10, 1 << 13), ((1 << 34) + 1, 35 << 29), (33 << 34, 35 << 34)]
KATS_PARAMS = [((64 * 4096, 64 * 4096), (43, 12288)), ((64 * 2**20, 64 * 2**20), (39, 7 * 2**19)),
               ((64 * 2**20, 64 * 2**18), (40, 2**21)), ((64 * 2**22, 64 * 2**22), (38, 7 * 2**21)),
               ((64 * 144000, 64 * 144000), (41, 7 << 16))]


def test_ntt_len_and_params_after_init():
    for T, want in KATS_LEN:
        assert engine.ntt_len(T) == want == O.ntt_len(T)
    for (ba, bb), want in KATS_PARAMS:
        assert engine.fp_ntt_params(ba, bb) == want == O.ntt_params(ba, bb)
    rng = np.random.default_rng(5)
    for _ in range(300):
        ba, bb = (int(v) for v in rng.integers(1, 64 * 2**22, 2))
        assert engine.fp_ntt_params(ba, bb) == O.ntt_params(ba, bb)
        T = int(rng.integers(1, 315 << 30))
        assert engine.ntt_len(T) == O.ntt_len(T)
    with pytest.raises(engine.ArgumentError):
        engine.ntt_len((315 << 34) + 1)


def test_magnitude_bits_read_on_the_device():
    """bits = -1 makes the library read each operand's top limb on the device;
    zero top limbs give a smaller magnitude (and a shorter transform) but the
    same product."""
    for m, n, ztop in [(3000, 2000, 0), (3000, 2000, 5), (5000, 700, 1000), (64, 64, 0)]:
        a, b = wl.operand("test", 60 + m, 0, m), wl.operand("test", 60 + m, 1, n)
        if ztop:
            a[-ztop:] = 0
        ba = engine.magnitude_bits(a, 0, m)
        assert ba == O.magnitude_bits(a)
        r = D.empty_limbs(m + n)
        D.mul_dev(r, D.to_dev(a), D.to_dev(b))          # bits read on the device
        torch.cuda.synchronize()
        assert np.array_equal(D.to_host(r), gmp_ref.mul(a, b))
