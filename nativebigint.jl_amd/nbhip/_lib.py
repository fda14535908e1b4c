"""ctypes binding of libnbh.so (the C ABI declared in include/nbh.h).

The library is the HIP build in this package directory.  There is no CPU
fallback: if the shared library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
import re

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libnbh.so")
HEADER = os.path.join(ROOT, "include", "nbh.h")

NBH_OK = 0
NBH_ERR_LENGTH = 1
NBH_ERR_ARG = 2
NBH_ERR_HIP = 3
NBH_ERR_OOM = 4
NBH_ERR_NOT_INIT = 5
NBH_ERR_BUDGET = 6


class NbhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nbh error {code}: {msg}")
        self.code = code


class ArgumentError(NbhError, ValueError):
    """The reference's ArgumentError: transform length beyond 315*2^34
    (src/fpntt.jl:948-949)."""


class nbh_job(ctypes.Structure):
    _fields_ = [("r", ctypes.c_void_p), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("m", ctypes.c_int64), ("n", ctypes.c_int64),
                ("bits_a", ctypes.c_int64), ("bits_b", ctypes.c_int64), ("k", ctypes.c_int64)]


class nbh_hgcd_job(ctypes.Structure):
    _fields_ = [("m00", ctypes.c_void_p), ("m01", ctypes.c_void_p), ("m10", ctypes.c_void_p),
                ("m11", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("mn", ctypes.c_int64), ("n", ctypes.c_int64),
                ("r0", ctypes.c_void_p), ("r1", ctypes.c_void_p),
                ("s0", ctypes.c_void_p), ("s1", ctypes.c_void_p),
                ("ra", ctypes.c_void_p), ("rb", ctypes.c_void_p), ("nn", ctypes.c_int64)]


class nbh_hgcd_level_job(ctypes.Structure):
    _fields_ = [("m00", ctypes.c_void_p), ("m01", ctypes.c_void_p), ("m10", ctypes.c_void_p),
                ("m11", ctypes.c_void_p), ("mn", ctypes.c_int64),
                ("n00", ctypes.c_void_p), ("n01", ctypes.c_void_p), ("n10", ctypes.c_void_p),
                ("n11", ctypes.c_void_p), ("m1n", ctypes.c_int64),
                ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("p", ctypes.c_int64),
                ("ra", ctypes.c_void_p), ("rb", ctypes.c_void_p), ("nn", ctypes.c_int64),
                ("sa", ctypes.c_void_p), ("sb", ctypes.c_void_p),
                ("n_ab", ctypes.c_void_p), ("n_m", ctypes.c_void_p)]


class nbh_spec_job(ctypes.Structure):
    _fields_ = [("r", ctypes.c_void_p), ("a", ctypes.c_void_p), ("m", ctypes.c_int64),
                ("bits_a", ctypes.c_int64), ("s", ctypes.c_void_p)]


class nbh_barrett_job(ctypes.Structure):
    _fields_ = [("t", ctypes.c_void_p), ("r", ctypes.c_void_p)]


NBH_HGCD_FORWARD = 0
NBH_HGCD_ADJUST = 1
NBH_HGCD_ADJUST_FULL = 2

_i32, _i64, _vp, _d = ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
_P = ctypes.POINTER

SIGNATURES = {
    "nbh_init": (_i32, [_i32]),
    "nbh_shutdown": (_i32, []),
    "nbh_last_error_string": (ctypes.c_char_p, []),
    "nbh_ntt_len": (_i32, [_i64, _P(_i64)]),
    "nbh_ntt_params": (_i32, [_i64, _i64, _P(_i32), _P(_i64)]),
    "nbh_mul": (_i32, [_vp, _vp, _i64, _vp, _i64]),
    "nbh_sqr": (_i32, [_vp, _vp, _i64]),
    "nbh_mul_dev": (_i32, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _vp]),
    "nbh_sqr_dev": (_i32, [_vp, _vp, _i64, _i64, _vp]),
    "nbh_mul_dev_noalloc": (_i32, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _vp]),
    "nbh_mullo": (_i32, [_vp, _vp, _i64, _vp, _i64, _i64]),
    "nbh_sqrlo": (_i32, [_vp, _vp, _i64, _i64]),
    "nbh_mullo_dev": (_i32, [_vp, _vp, _i64, _vp, _i64, _i64, _i64, _i64, _vp]),
    "nbh_sqrlo_dev": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp]),
    "nbh_batch_create": (_i32, [_P(nbh_job), _i64, _P(_vp)]),
    "nbh_batch_run": (_i32, [_vp, _vp]),
    "nbh_batch_destroy": (None, [_vp]),
    "nbh_mul_batch_dev": (_i32, [_P(nbh_job), _i64, _vp]),
    "nbh_mul_batch": (_i32, [_P(nbh_job), _i64]),
    "nbh_server_create": (_i32, [_i64, _i32, _i32, _P(_vp)]),
    "nbh_server_submit": (_i32, [_vp, _P(nbh_job), _P(_i64)]),
    "nbh_server_poll": (_i32, [_vp, _P(_i64), _i64, _P(_i64)]),
    "nbh_server_drain": (_i32, [_vp]),
    "nbh_server_stats": (_i32, [_vp, _P(_i64), _P(_i64), _P(_i64)]),
    "nbh_server_destroy": (None, [_vp]),
    "nbh_alloc": (_i32, [_i64, _P(_vp)]),
    "nbh_free": (None, [_vp]),
    "nbh_upload": (_i32, [_vp, _vp, _i64]),
    "nbh_download": (_i32, [_vp, _vp, _i64]),
    "nbh_batch_arena_bytes": (_i32, [_P(nbh_job), _i64, _P(_i64)]),
    "nbh_batch_create_in": (_i32, [_P(nbh_job), _i64, _vp, _i64, _vp, _P(_vp)]),
    "nbh_batch_release": (None, [_vp]),
    "nbh_hgcd_create": (_i32, [_P(nbh_hgcd_job), _i64, _i32, _P(_vp)]),
    "nbh_hgcd_run": (_i32, [_vp, _vp]),
    "nbh_hgcd_destroy": (None, [_vp]),
    "nbh_hgcd_level_create": (_i32, [_P(nbh_hgcd_level_job), _i64, _P(_vp)]),
    "nbh_spectrum_create": (_i32, [_vp, _i64, _i64, _i64, _vp, _P(_vp)]),
    "nbh_spectrum_destroy": (None, [_vp]),
    "nbh_spectrum_info": (_i32, [_vp, _P(_i64), _P(_i32), _P(_i64), _P(_i64)]),
    "nbh_mul_spec_dev": (_i32, [_vp, _vp, _i64, _i64, _vp, _vp]),
    "nbh_spec_batch_create": (_i32, [_P(nbh_spec_job), _i64, _P(_vp)]),
    "nbh_set_streams": (_i32, [_i32]),
    "nbh_batch_set_streams": (_i32, [_vp, _i32]),
    "nbh_batch_capture": (_i32, [_vp]),
    "nbh_batch_replay": (_i32, [_vp, _vp]),
    "nbh_barrett_create": (_i32, [_vp, _i64, _vp, _i64, _vp, _P(_vp)]),
    "nbh_barrett_destroy": (None, [_vp]),
    "nbh_barrett_batch_create": (_i32, [_vp, _P(nbh_barrett_job), _i64, _P(_vp)]),
    "nbh_barrett_batch_run": (_i32, [_vp, _vp]),
    "nbh_barrett_batch_destroy": (None, [_vp]),
    "nbh_ntt_fwd_dev": (_i32, [_vp, _i64, _i32, _vp]),
    "nbh_ntt_rev_dev": (_i32, [_vp, _i64, _i32, _vp]),
    "nbh_pack_dev": (_i32, [_vp, _i64, _vp, _i64, _i32, _i64, _vp]),
    "nbh_pointwise_dev": (_i32, [_vp, _vp, _i64, _i32, _vp]),
    "nbh_unpack_dev": (_i32, [_vp, _i64, _vp, _vp, _i64, _i64, _i32, _vp]),
    "nbh_plan_nstages": (_i32, [_i64, _i32, _P(_i32)]),
    "nbh_plan_stage": (_i32, [_i64, _i32, _i32, _P(_i32), _P(_i64), _P(_i64), _vp]),
    "nbh_plan_twiddle_row": (_i32, [_i64, _i32, _i32, _i32, _vp]),
    "nbh_plan_ninv": (_i32, [_i64, _i32, _P(_d), _P(_d)]),
    "nbh_set_tuning": (_i32, [_i32, _i32, _i32, _i32]),
    "nbh_get_tuning": (_i32, [_P(_i32), _P(_i32), _P(_i32), _P(_i32)]),
    "nbh_plan_groups": (_i32, [_i64, _P(_i32), _vp]),
    "nbh_set_tuning_table": (_i32, [_vp, _vp, _i32]),
    "nbh_set_tuning_table_single": (_i32, [_vp, _vp, _i32]),
    "nbh_get_tuning_for_single": (_i32, [_i64, _vp]),
    "nbh_get_tuning_for": (_i32, [_i64, _vp]),
    "nbh_profile_enable": (_i32, [_i32]),
    "nbh_profile_read": (_i32, [_vp, _vp, _i32]),
    "nbh_synchronize": (_i32, [_vp]),
    "nbh_device_info": (_i32, [_P(_i32), _P(_i64), _P(_i64), ctypes.c_char_p, _i32]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function name declared in include/nbh.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(nbh_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                               "(there is no CPU fallback for the HIP engine)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc == NBH_OK:
        return
    msg = (lib().nbh_last_error_string() or b"").decode()
    if rc == NBH_ERR_LENGTH:
        raise ArgumentError(rc, msg)
    raise NbhError(rc, msg)


_initialised = False


def init(device: int = 0) -> None:
    global _initialised
    if not _initialised:
        check(lib().nbh_init(device))
        _initialised = True


def shutdown() -> None:
    """nbh_shutdown: raises NbhError (NBH_ERR_ARG) while handles are live."""
    global _initialised
    check(lib().nbh_shutdown())
    _initialised = False
