"""CPU checks of the C ABI: the library loads, exports every symbol declared
in include/nbh.h, and its host-side sizing matches the oracle.  No compute
call is made (there is no GPU here)."""
import ctypes

import pytest

from oracle import oracle as O
from tests import helpers  # noqa: F401  (puts the package on sys.path)
from nbhip import _lib, engine


def test_library_exports_every_header_symbol():
    L = _lib.lib()
    syms = _lib.header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []
    assert set(syms) == set(_lib.SIGNATURES)


@pytest.mark.parametrize("T", [1, 4, 5, 100, 1023, 1025, 13000, 15360, 20000, (15 << 14) + 1, 1 << 20,
                               (7 << 22) - 1, 1 << 25, (1 << 34) + 1, 33 << 34, 53 << 34, 315 << 34])
def test_ntt_len_matches_oracle(T):
    assert engine.ntt_len(T) == O.ntt_len(T)


def test_ntt_len_argument_error():
    with pytest.raises(engine.ArgumentError):
        engine.ntt_len((315 << 34) + 1)
    with pytest.raises(ValueError):          # ArgumentError is a ValueError
        engine.ntt_len((315 << 34) + 1)


@pytest.mark.parametrize("ba,bb", [(64, 64), (64 * 5, 64 * 3), (64 * 152, 64 * 152), (64 * 4096, 64 * 4096),
                                   (64 * 2**20, 64 * 2**20), (64 * 2**20, 64 * 2**18), (64 * 2**22, 64 * 2**22),
                                   (9000, 13), (64 * 144000, 64 * 144000), (1, 1)])
def test_params_match_oracle(ba, bb):
    assert engine.fp_ntt_params(ba, bb) == O.ntt_params(ba, bb)


def test_entry_points_refuse_before_init():
    L = _lib.lib()
    assert L.nbh_mul(None, None, 1, None, 1) == _lib.NBH_ERR_NOT_INIT
    assert b"nbh_init" in L.nbh_last_error_string()
    n = ctypes.c_int32()
    assert L.nbh_plan_nstages(16, 0, ctypes.byref(n)) == _lib.NBH_ERR_NOT_INIT


def test_host_mirror_argument_checks():
    import numpy as np
    a = np.ones(200, dtype=np.uint64)
    with pytest.raises(NotImplementedError):
        engine.mul(np.empty(100, dtype=np.uint64), 0, a, 0, 50, a, 0, 50)
    with pytest.raises(ValueError):
        engine.mul_fpntt2(a, 0, a, 0, 100, a, 100, 100)       # r aliases a
    with pytest.raises(TypeError):
        engine.mul_fpntt2(np.empty(10), 0, a, 0, 5, a, 0, 5)


def test_newer_entry_points_refuse_before_init():
    """Every entry point that needs the device answers NBH_ERR_NOT_INIT (with
    a message) instead of touching it when nbh_init has not succeeded."""
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.nbh_spectrum_create(None, 1, -1, 64, None, ctypes.byref(h)) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_barrett_create(None, 1, None, 2, None, ctypes.byref(h)) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_batch_capture(None) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_batch_replay(None, None) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_set_streams(0) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_batch_set_streams(None, 0) == _lib.NBH_ERR_NOT_INIT
    assert L.nbh_mul_spec_dev(None, None, 1, -1, None, None) == _lib.NBH_ERR_NOT_INIT
    assert b"nbh_init" in L.nbh_last_error_string()
