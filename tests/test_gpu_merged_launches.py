"""Batches whose mid tiles share a block size but not an LDS size: whole
transforms of 512..1024 points (64 threads each) and of 1152..2048 points
(128 threads each) run as one mid launch per block size, sized to the
largest tile's LDS; and tiles of different block sizes with the same blocks
per CU share the launch of the largest (its extra threads idle, general
passes instead of edge passes): 6144 and 8192 points (384 and 512 threads,
two blocks per CU) and 9216..16384 points (576..1024 threads, one block per
CU) run as two launches.  Every product and square stays bit-exact against
libgmp (uniform, all-ones and 2^k +- 1 operands); the per-kernel profile
counts the launches."""
import numpy as np
import pytest

from tests import gmp_ref
from tests.helpers import rand_big_style

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402

# balanced sizes (limbs) whose transform lengths are 512, 576, 640, 768, 1024
# (one wave per tile) and 1152, 1280, 1536, 2048 (two waves per tile)
SIZES = [152, 184, 216, 232, 280, 376, 408, 456, 552]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()
    engine.init()


def _mid_threads(N: int) -> int:
    # one thread per 16 points, whole waves, 64..1024
    t = (N // 16 + 63) // 64 * 64
    return min(max(t, 64), 1024)


@pytest.mark.parametrize("squares", [False, True])
def test_one_group_tiles_share_launches(squares):
    rng = np.random.default_rng(2024 + squares)
    lens = {}
    for m in SIZES:
        _, N = engine.fp_ntt_params(64 * m, 64 * m)
        lens[m] = N
        assert len(D.plan_groups(N)) == 1, (m, N, D.plan_groups(N))
    assert len(set(lens.values())) == len(SIZES)
    jobs, want = [], []
    for rep in range(4):
        for m in SIZES:
            style = 1 + (rep % 3)
            a = rand_big_style(rng, m, style)
            if squares and rep % 2:
                want.append(gmp_ref.sqr(a))
                jobs.append((D.empty_limbs(2 * m), D.to_dev(a), None))
            else:
                b = rand_big_style(rng, m, 1 + ((rep + 1) % 3))
                want.append(gmp_ref.mul(a, b))
                jobs.append((D.empty_limbs(2 * m), D.to_dev(a), D.to_dev(b)))
    B = D.Batch(jobs)
    D.profile_enable(True)
    try:
        D.profile_read()
        B.run()
        torch.cuda.synchronize()
        prof = D.profile_read()
    finally:
        D.profile_enable(False)
        B.close()
    for i, (j, w) in enumerate(zip(jobs, want)):
        assert np.array_equal(D.to_host(j[0]), w), (i, j[1].numel())
    if not squares:
        # one mid launch per block size instead of one per transform length
        assert prof["mid"][1] == len({_mid_threads(N) for N in lens.values()}), prof
    else:
        assert prof["mid"][1] < len(SIZES), prof


# balanced sizes whose transforms are 6144, 8192 (two blocks per CU) and
# 9216, 10240, 12288, 16384 points (one block per CU), one group each
SIZES_WIDE = {6144: 2000, 8192: 2120, 9216: 2808, 10240: 3104, 12288: 3448, 16384: 4136}


def test_tiles_join_launches_of_larger_blocks():
    rng = np.random.default_rng(77)
    for N, m in SIZES_WIDE.items():
        assert engine.fp_ntt_params(64 * m, 64 * m)[1] == N, (m, N)
        assert len(D.plan_groups(N)) == 1, (m, N, D.plan_groups(N))
    jobs, want = [], []
    for rep in range(3):
        for m in SIZES_WIDE.values():
            a = rand_big_style(rng, m, 1 + rep)
            if rep == 2:
                want.append(gmp_ref.sqr(a))
                jobs.append((D.empty_limbs(2 * m), D.to_dev(a), None))
            else:
                b = rand_big_style(rng, m, 3 - rep)
                want.append(gmp_ref.mul(a, b))
                jobs.append((D.empty_limbs(2 * m), D.to_dev(a), D.to_dev(b)))
    B = D.Batch(jobs)
    D.profile_enable(True)
    try:
        D.profile_read()
        B.run()
        torch.cuda.synchronize()
        prof = D.profile_read()
    finally:
        D.profile_enable(False)
        B.close()
    for i, (j, w) in enumerate(zip(jobs, want)):
        assert np.array_equal(D.to_host(j[0]), w), (i, j[1].numel())
    # 384/512-thread tiles in one launch, 576..1024-thread tiles in another
    assert prof["mid"][1] == 2, prof
