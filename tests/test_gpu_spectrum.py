"""Products against cached forward spectra of a fixed operand (the modulus
and reciprocal that barrett_reduce!, src/barrett.jl:44-81, multiplies by on
every call): bit-exact against Python big ints and against the plain device
product, alone and in batches, at single- and multi-group lengths."""
import numpy as np
import pytest

from tests.helpers import ints, rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402
from nbhip._lib import NbhError  # noqa: E402


def _limbs(rng, n, style=1):
    return rand_big_style(rng, n, style)


# (limbs of b, limbs of the longest a, limbs of the a's tried)
CASES = [
    (160, 160, [160, 100, 1]),          # shortest engine products, one group
    (700, 700, [700, 699, 350]),        # Barrett-like: k x k and k x (k+1)
    (3000, 3001, [3001, 2000, 17]),     # odd-length family
    (9000, 9000, [9000, 4500]),         # several groups
    (40000, 20000, [20000, 19999]),     # unbalanced, several groups
]


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init()


@pytest.mark.parametrize("nb,amax,alens", CASES)
def test_spectrum_products(nb, amax, alens):
    rng = np.random.default_rng(nb)
    b = _limbs(rng, nb)
    db = D.to_dev(b)
    S = D.Spectrum(db, 64 * amax)
    assert S.n == nb and S.bits_a_max == 64 * amax
    B = ints(b)
    for m in alens:
        for style in (1, 2):
            a = _limbs(rng, m, style)
            da = D.to_dev(a)
            r = D.empty_limbs(m + nb)
            S.mul(r, da)
            r2 = D.empty_limbs(m + nb)
            D.mul_dev(r2, da, db)
            torch.cuda.synchronize()
            got = D.to_host(r)
            assert np.array_equal(got, D.to_host(r2)), (nb, m, style)
            if m * nb <= 3000 * 3001:
                assert ints(got) == ints(a) * B, (nb, m, style)
    S.close()


def test_all_ones_at_the_exactness_bound():
    """All-ones operands maximise every chunk and so every convolution
    coefficient: the spectrum's chunk width must stay exact for the longest a."""
    rng = np.random.default_rng(7)
    for nb, amax in ((400, 400), (5000, 2000)):
        b = _limbs(rng, nb, 2)
        a = _limbs(rng, amax, 2)
        S = D.Spectrum(D.to_dev(b), 64 * amax)
        r = D.empty_limbs(amax + nb)
        S.mul(r, D.to_dev(a))
        torch.cuda.synchronize()
        assert ints(D.to_host(r)) == ints(a) * ints(b)
        S.close()


def test_spectrum_reused_and_unchanged():
    """Many products against one spectrum, then the first again: the spectrum
    is only read."""
    rng = np.random.default_rng(11)
    nb = 2048
    b = _limbs(rng, nb)
    S = D.Spectrum(D.to_dev(b), 64 * 2100)
    outs = []
    As = [_limbs(rng, m) for m in (2100, 1500, 2048, 900)]
    for a in As:
        r = D.empty_limbs(a.size + nb)
        S.mul(r, D.to_dev(a))
        outs.append(r)
    r0 = D.empty_limbs(As[0].size + nb)
    S.mul(r0, D.to_dev(As[0]))
    torch.cuda.synchronize()
    for a, r in zip(As, outs):
        assert ints(D.to_host(r)) == ints(a) * ints(b)
    assert np.array_equal(D.to_host(r0), D.to_host(outs[0]))
    S.close()


def test_longer_operand_is_rejected():
    rng = np.random.default_rng(3)
    b = _limbs(rng, 300)
    S = D.Spectrum(D.to_dev(b), 64 * 300)
    a = _limbs(rng, 301)
    with pytest.raises(NbhError):
        S.mul(D.empty_limbs(601), D.to_dev(a))
    # more limbs but a zero top limb: magnitude_bits 64*300 fits the bound
    a[-1] = 0
    r = D.empty_limbs(601)
    S.mul(r, D.to_dev(a))
    torch.cuda.synchronize()
    assert ints(D.to_host(r)) == ints(a) * ints(b)
    S.close()


def test_zero_operands():
    rng = np.random.default_rng(5)
    zb = D.to_dev(np.zeros(200, dtype=np.uint64))
    S0 = D.Spectrum(zb, 64 * 200)
    r = D.to_dev(np.full(400, 7, dtype=np.uint64))
    S0.mul(r, D.to_dev(_limbs(rng, 200)))
    S = D.Spectrum(D.to_dev(_limbs(rng, 200)), 64 * 200)
    r2 = D.to_dev(np.full(400, 7, dtype=np.uint64))
    S.mul(r2, D.to_dev(np.zeros(200, dtype=np.uint64)))
    torch.cuda.synchronize()
    assert not D.to_host(r).any() and not D.to_host(r2).any()
    S0.close()
    S.close()


def test_spectrum_batch_equals_alone():
    """A batch mixing two spectra and several a lengths gives the same limbs
    as each product run alone."""
    rng = np.random.default_rng(13)
    specs = []
    for nb, amax in ((640, 641), (12000, 6000)):
        b = _limbs(rng, nb)
        specs.append((b, D.Spectrum(D.to_dev(b), 64 * amax), amax))
    jobs, host = [], []
    for i in range(24):
        b, S, amax = specs[i % 2]
        m = int(rng.integers(max(1, amax // 3), amax + 1))
        a = _limbs(rng, m, 1 + i % 4)
        da = D.to_dev(a)
        r = D.empty_limbs(m + b.size)
        jobs.append((r, da, S))
        host.append((a, b, S, da))
    SB = D.SpecBatch(jobs)
    for rep in range(2):
        for r, _, _ in jobs:
            r.fill_(-1)
        SB.run()
        torch.cuda.synchronize()
        for (r, _, _), (a, b, S, da) in zip(jobs, host):
            alone = D.empty_limbs(a.size + b.size)
            S.mul(alone, da)
            torch.cuda.synchronize()
            got = D.to_host(r)
            assert np.array_equal(got, D.to_host(alone))
            if a.size * b.size < 10_000_000:
                assert ints(got) == ints(a) * ints(b)
    SB.close()
    for _, S, _ in specs:
        S.close()
