"""The oracle against the committed libgmp fixtures (CPU)."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import wl

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "products.json")))


def sha(r):
    return hashlib.sha256(np.ascontiguousarray(r, dtype="<u8").tobytes()).hexdigest()


def _product(key, fn_mul, fn_sqr):
    wname, job, shape = key.split("/")[:3]
    m, n = (int(v) for v in shape.split("x"))
    a = wl.operand(wname, int(job), 0, m)
    if key.endswith("/sqr"):
        return fn_sqr(a)
    return fn_mul(a, wl.operand(wname, int(job), 1, n))


@pytest.mark.parametrize("key", [k for k in FIX["products"] if not k.startswith("large/")])
def test_oracle_matches_fixture(key):
    assert sha(_product(key, O.mul, O.sqr)) == FIX["products"][key]


@pytest.mark.parametrize("key", list(FIX["small"]))
def test_oracle_matches_small_vectors(key):
    want = np.array([int(h, 16) for h in FIX["small"][key]], dtype=np.uint64)
    assert np.array_equal(_product(key, O.mul, O.sqr), want)
