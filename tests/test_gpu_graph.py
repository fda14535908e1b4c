"""A planned batch captured into a HIP graph: every replay gives the eager
run's limbs, with other eager device work (fills, another batch, torch ops)
between the replays, on the capture stream and on another stream, and with
the launches forked onto side streams or issued in order."""
import numpy as np
import pytest

from tests.helpers import wl

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
from nbhip import device as D  # noqa: E402
from nbhip import engine  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _init():
    engine.init()
    yield
    D.set_streams(8)


def _batch(count, seed):
    jobs = wl.batch_jobs(count=count, lo=1, hi=12000, workload="test")
    ops = wl.batch_operands(jobs, workload="test")
    tens = []
    for j, (a, b) in zip(jobs[seed:], ops[seed:]):
        r = D.empty_limbs(j.m + (j.m if b is None else j.n))
        tens.append((r, D.to_dev(a), None if b is None else D.to_dev(b)))
    return tens


@pytest.mark.parametrize("nside", [8, 4, 0])
def test_graph_replays_equal_eager(nside):
    D.set_streams(nside)
    tens = _batch(80, 0)
    other = _batch(40, 7)
    B = D.Batch(tens)
    B2 = D.Batch(other)
    B.run()
    torch.cuda.synchronize()
    want = [D.to_host(r) for r, _, _ in tens]
    B.capture()
    s2 = torch.cuda.Stream()
    for rep in range(4):
        # eager work between replays: clobber the outputs, run another batch,
        # and a torch reduction over many blocks
        for r, _, _ in tens:
            r.fill_(-1)
        B2.run()
        junk = torch.ones(1 << 20, device="cuda").sum()
        stream = s2 if rep % 2 else torch.cuda.current_stream()
        if rep % 2:
            s2.wait_stream(torch.cuda.current_stream())
        B.replay(stream)
        torch.cuda.synchronize()
        assert float(junk) == float(1 << 20)
        for (r, _, _), w in zip(tens, want):
            assert np.array_equal(D.to_host(r), w), (nside, rep)
    # eager runs still work after capture
    for r, _, _ in tens:
        r.fill_(0)
    B.run()
    torch.cuda.synchronize()
    for (r, _, _), w in zip(tens, want):
        assert np.array_equal(D.to_host(r), w)
    B.close()
    B2.close()


def test_stream_settings_are_checked():
    from nbhip._lib import NbhError, check, lib
    with pytest.raises(NbhError):
        D.set_streams(9)
    with pytest.raises(NbhError):
        D.set_streams(-1)
    tens = _batch(4, 0)
    B = D.Batch(tens)
    with pytest.raises(NbhError):
        check(lib().nbh_batch_set_streams(B._h, 9))
    check(lib().nbh_batch_set_streams(B._h, -1))
    B.run()
    torch.cuda.synchronize()
    B.close()
    D.set_streams(8)
