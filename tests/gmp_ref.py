"""libgmp's mpn_mul / mpn_sqr through ctypes: the exact-product reference.

The integer product is unique, so bit-exact against GMP is bit-exact against
the reference's Base.BigInt differential tests (test/test_fpntt.jl:288-331).
Returns None from `available()` when libgmp is absent.
"""
import ctypes
import ctypes.util
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

_gmp = None


def _load():
    global _gmp
    if _gmp is None:
        for name in ("libgmp.so.10", ctypes.util.find_library("gmp")):
            if not name:
                continue
            try:
                _gmp = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _gmp is not None:
            _gmp.__gmpn_mul.restype = ctypes.c_uint64
            _gmp.__gmpn_mul.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                        ctypes.c_void_p, ctypes.c_long]
            _gmp.__gmpn_sqr.restype = None
            _gmp.__gmpn_sqr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    return _gmp


def available() -> bool:
    return _load() is not None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    if a.size < b.size:
        a, b = b, a
    r = np.empty(a.size + b.size, dtype=np.uint64)
    _load().__gmpn_mul(_p(r), _p(a), a.size, _p(b), b.size)
    return r


def sqr(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    r = np.empty(2 * a.size, dtype=np.uint64)
    _load().__gmpn_sqr(_p(r), _p(a), a.size)
    return r


def products(pairs, threads: int | None = None) -> list:
    """libgmp products of many (a, b) pairs (b None: a square) on a thread
    pool; ctypes releases the GIL during each call."""
    threads = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda ab: sqr(ab[0]) if ab[1] is None else mul(ab[0], ab[1]), pairs))
