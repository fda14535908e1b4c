"""Host-side mirror of NativeBigInt.jl's NTT multiplication interface.

Same names, argument meaning and error behaviour as the reference's engine
(src/fpntt.jl:982-1037) and the dispatch lines that reach it (src/mul.jl:
24-26, 127-164, 383-389), over numpy uint64 limb buffers with 0-based
offsets (the reference's `Memory{UInt64}` plus offset).  Every product runs on
the GPU through libnbh.so; below-threshold sizes, which the reference sends to
Karatsuba/basecase, are outside this engine and raise NotImplementedError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import ArgumentError, NbhError, check, init, lib

# src/mul.jl:24-26
MUL_FPNTT_MIN = 64
MUL_FPNTT_THRESHOLD = 152
SQR_FPNTT_THRESHOLD = 160

__all__ = ["MUL_FPNTT_MIN", "MUL_FPNTT_THRESHOLD", "SQR_FPNTT_THRESHOLD", "ArgumentError", "NbhError",
           "magnitude_bits", "ntt_len", "fp_ntt_params", "mul_fpntt2", "sqr_fpntt2", "mul", "sqr", "mul_bal",
           "mullo", "sqrlo", "mul_batch", "mul_batch_records", "init"]


def _u64(buf) -> np.ndarray:
    if not isinstance(buf, np.ndarray) or buf.dtype != np.uint64 or not buf.flags.c_contiguous:
        raise TypeError("limb buffers are C-contiguous numpy uint64 arrays")
    return buf


def _ptr(buf: np.ndarray, off: int) -> int:
    return buf.ctypes.data + 8 * off


def magnitude_bits(a, ao: int, n: int) -> int:
    """src/kernels/addsub.jl:150: 64(n-1) + bit length of the top limb."""
    return 64 * (n - 1) + int(a[ao + n - 1]).bit_length()


def ntt_len(T: int) -> int:
    """src/fpntt.jl:936-951; raises ArgumentError above 315*2^34."""
    N = ctypes.c_int64()
    check(lib().nbh_ntt_len(T, ctypes.byref(N)))
    return N.value


def fp_ntt_params(bits_a: int, bits_b: int):
    """src/fpntt.jl:964-977 -> (b, N)."""
    b = ctypes.c_int32()
    N = ctypes.c_int64()
    check(lib().nbh_ntt_params(bits_a, bits_b, ctypes.byref(b), ctypes.byref(N)))
    return b.value, N.value


def _check_ranges(r, ro, rlen, a, ao, m, b=None, bo=0, n=0):
    if ro < 0 or ro + rlen > r.size or ao < 0 or ao + m > a.size or m < 1:
        raise ValueError("limb range out of bounds")
    if b is not None and (bo < 0 or bo + n > b.size or n < 1):
        raise ValueError("limb range out of bounds")
    rv = r[ro:ro + rlen]
    if np.may_share_memory(rv, a[ao:ao + m]) or (b is not None and np.may_share_memory(rv, b[bo:bo + n])):
        raise ValueError("r must not alias a or b (src/mul.jl:380-381)")


def mul_fpntt2(r, ro: int, a, ao: int, m: int, b, bo: int, n: int) -> None:
    """r[ro:ro+m+n] = a[ao:ao+m] * b[bo:bo+n] (src/fpntt.jl:982-1014)."""
    r, a, b = _u64(r), _u64(a), _u64(b)
    _check_ranges(r, ro, m + n, a, ao, m, b, bo, n)
    init()
    check(lib().nbh_mul(_ptr(r, ro), _ptr(a, ao), m, _ptr(b, bo), n))


def sqr_fpntt2(r, ro: int, a, ao: int, n: int) -> None:
    """r[ro:ro+2n] = a[ao:ao+n]^2 (src/fpntt.jl:1016-1037)."""
    r, a = _u64(r), _u64(a)
    _check_ranges(r, ro, 2 * n, a, ao, n)
    init()
    check(lib().nbh_sqr(_ptr(r, ro), _ptr(a, ao), n))


def mul(r, ro: int, a, ao: int, m: int, b, bo: int, n: int) -> None:
    """mul! (src/mul.jl:383-389), m >= n >= 1: the NTT branch."""
    if m < n:
        raise ValueError("mul! requires m >= n (callers swap)")
    if not (n >= MUL_FPNTT_MIN and m + n >= 2 * MUL_FPNTT_THRESHOLD):
        raise NotImplementedError("below the NTT thresholds the reference uses Karatsuba/basecase")
    mul_fpntt2(r, ro, a, ao, m, b, bo, n)


def mul_bal(r, ro: int, a, ao: int, b, bo: int, n: int) -> None:
    """mul_bal! (src/mul.jl:154-164): balanced n x n, NTT branch at n >= 152."""
    if n < MUL_FPNTT_THRESHOLD:
        raise NotImplementedError("below the NTT threshold the reference uses Karatsuba/basecase")
    mul_fpntt2(r, ro, a, ao, n, b, bo, n)


def sqr(r, ro: int, a, ao: int, n: int) -> None:
    """sqr! (src/mul.jl:141-150): NTT branch at n >= 160."""
    if n < SQR_FPNTT_THRESHOLD:
        raise NotImplementedError("below the NTT threshold the reference uses Karatsuba/basecase")
    sqr_fpntt2(r, ro, a, ao, n)


def mullo(r, ro: int, a, ao: int, la: int, b, bo: int, lb: int, k: int) -> None:
    """mullo! (src/mul.jl:305-354): r[ro:ro+k] = a*b mod beta^k, la, lb,
    k >= 1, r not aliasing a or b.  Operand limbs at or above k are ignored,
    as the reference clamps la and lb to k.  Exactly k limbs are written: the
    reference's two clobber limbs above k are not needed.  Every size runs on
    the GPU (schoolbook below the mul! NTT thresholds)."""
    r, a, b = _u64(r), _u64(a), _u64(b)
    if k < 1:
        raise ValueError("k >= 1")
    la, lb = min(la, k), min(lb, k)
    _check_ranges(r, ro, k, a, ao, la, b, bo, lb)
    init()
    check(lib().nbh_mullo(_ptr(r, ro), _ptr(a, ao), la, _ptr(b, bo), lb, k))


def sqrlo(r, ro: int, a, ao: int, la: int, k: int) -> None:
    """sqrlo! (src/mul.jl:360-378): r[ro:ro+k] = a^2 mod beta^k."""
    r, a = _u64(r), _u64(a)
    if k < 1:
        raise ValueError("k >= 1")
    la = min(la, k)
    _check_ranges(r, ro, k, a, ao, la)
    init()
    check(lib().nbh_sqrlo(_ptr(r, ro), _ptr(a, ao), la, k))


def mul_batch_records(jobs):
    """The checked nbh_job records of a host-buffer batch (None if empty),
    for callers that issue the same batch repeatedly (nbh_mul_batch): each
    job is one mul!/sqr!/mullo!/sqrlo! of the reference (src/mul.jl:127-164,
    305-389).  A job is the tuple
    (r, ro, a, ao, m, b, bo, n, k): b None for a square (bo, n ignored), k 0
    for the full product r[ro:ro+m+n] (2m for a square), k >= 1 for the low
    short product r[ro:ro+k] with the operands clamped to k limbs.  Every size
    runs on the GPU (schoolbook below the mul!/sqr! NTT thresholds).  No
    result may overlap another job's result or any operand."""
    recs = []
    spans = []   # (start, end, is_result) byte ranges
    for j, job in enumerate(jobs):
        r, ro, a, ao, m, b, bo, n, k = job
        r, a = _u64(r), _u64(a)
        sqr = b is None
        if k < 0:
            raise ValueError(f"job {j}: k >= 0")
        if not sqr:
            b = _u64(b)
        mm = min(m, k) if k else m
        nn = mm if sqr else (min(n, k) if k else n)
        rl = k if k else mm + nn
        if sqr:
            _check_ranges(r, ro, rl, a, ao, mm)
        else:
            _check_ranges(r, ro, rl, a, ao, mm, b, bo, nn)
        pa = _ptr(a, ao)
        pb = 0 if sqr else _ptr(b, bo)
        pr = _ptr(r, ro)
        recs.append((pr, pa, pb, mm, nn, 0, 0, k))
        spans.append((pa, pa + 8 * mm, False))
        if not sqr:
            spans.append((pb, pb + 8 * nn, False))
        spans.append((pr, pr + 8 * rl, True))
    if not recs:
        return None
    spans.sort()
    end_all = end_res = 0
    for s, e, is_res in spans:
        if s < end_res or (is_res and s < end_all):
            raise ValueError("a result overlaps another job's result or an operand")
        end_all = max(end_all, e)
        if is_res:
            end_res = max(end_res, e)
    arr = (_lib.nbh_job * len(recs))()
    for i, (pr, pa, pb, mm, nn, _, _, k) in enumerate(recs):
        arr[i] = _lib.nbh_job(pr, pa, pb or None, mm, nn, 0, 0, k)
    return arr


def mul_batch(jobs) -> None:
    """Many independent products on host limb buffers in one call
    (nbh_mul_batch); see mul_batch_records for the job tuples."""
    arr = mul_batch_records(jobs)
    if arr is None:
        return
    init()
    check(lib().nbh_mul_batch(arr, len(arr)))
