"""Argument checks of the host-buffer batch (engine.mul_batch) that run
before any device call: results overlapping other results or operands are
rejected (r must not alias a or b, src/mul.jl:380-381; batch jobs are
independent), operands may be shared."""
import numpy as np
import pytest

from tests import helpers  # noqa: F401  (puts the package on sys.path)
from nbhip import engine


def test_result_overlaps_are_rejected():
    a = np.ones(300, dtype=np.uint64)
    r = np.zeros(1000, dtype=np.uint64)
    # two results overlapping by one limb
    with pytest.raises(ValueError, match="overlaps"):
        engine.mul_batch([(r, 0, a, 0, 200, a, 0, 200, 0), (r, 399, a, 0, 100, None, 0, 0, 0)])
    # a result over another job's operand
    with pytest.raises(ValueError, match="overlaps"):
        engine.mul_batch([(r, 0, a, 0, 100, None, 0, 0, 0), (r, 500, r, 150, 100, None, 0, 0, 0)])
    # a result over its own operand
    with pytest.raises(ValueError, match="alias"):
        engine.mul_batch([(r, 0, r, 100, 200, a, 0, 200, 0)])
    # a short product's result is k limbs: a neighbour just above it is fine
    # up to the device call, which this test does not reach
    with pytest.raises(ValueError, match="k >= 0"):
        engine.mul_batch([(r, 0, a, 0, 10, a, 0, 10, 5), (r, 5, a, 0, 10, None, 0, 0, -1)])


def test_bad_buffers_are_rejected():
    a = np.ones(10, dtype=np.uint64)
    r = np.zeros(10, dtype=np.uint64)
    with pytest.raises(TypeError):
        engine.mul_batch([(r, 0, a.astype(np.int64), 0, 5, None, 0, 0, 0)])
    with pytest.raises(ValueError, match="out of bounds"):
        engine.mul_batch([(r, 0, a, 0, 10, a, 0, 10, 0)])
