"""bench.py's per-kernel FP64 model splits the whole-product count exactly:
the forward groups, the mid kernel, the inverse groups and the unpack add up
to fp64_flops for products and squares, over stage lists of every family."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(spec)
spec.loader.exec_module(bench)

# (N, forward stage radices, groups as (La, Lb, S, C)); S is the radix product of each group
PLANS = [
    (12288, [3, 4, 4, 4, 4, 4, 4], [(12288, 1024, 12, 8), (1024, 1, 1024, 1)]),
    (3670016, [7, 4, 4, 4, 4, 4, 4, 4, 4, 4, 2],
     [(3670016, 8192, 448, 8), (8192, 1, 8192, 1)]),
    (2097152, [4] * 10 + [2], [(2097152, 32768, 64, 64), (32768, 2048, 16, 64), (2048, 1, 2048, 1)]),
    (512, [4, 4, 4, 4, 2], [(512, 1, 512, 1)]),
]


@pytest.mark.parametrize("N,radices,groups", PLANS)
@pytest.mark.parametrize("sqr", [False, True])
def test_split_adds_up(N, radices, groups, sqr):
    nconv = N - 7
    parts = bench.fp64_by_kind(groups, radices, N, sqr, nconv)
    assert sum(parts.values()) == pytest.approx(bench.fp64_flops(N, radices, sqr, nconv), rel=1e-12)
    if len(groups) == 1:
        assert parts["group_fwd"] == 0 and parts["group_inv"] == 0
    else:
        # forward groups transform every operand, inverse groups the product only
        assert parts["group_fwd"] == pytest.approx((1 if sqr else 2) * parts["group_inv"])
