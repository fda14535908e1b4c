"""ctypes front end to the CPU restatement of src/fpntt.jl (fpntt_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker (or the timed
CPU baseline) -- never as the thing measured or shipped.  The product library
(nativebigint.jl_amd) neither imports nor links it.

Each wrapper names the reference function it restates (file:line in the
reference repository).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_V3 = os.path.join(_HERE, "libfpntt_oracle.so")       # AVX2 + FMA (x86-64-v3)
_LIB_V4 = os.path.join(_HERE, "libfpntt_oracle_v4.so")    # AVX-512 (x86-64-v4), 8 doubles a vector

_lib = None
_variant = None


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512dq", "avx512bw", "avx512vl"} <= fl
    except OSError:
        pass
    return False


_forced = None


def _lib_path() -> str:
    """The AVX-512 build where the host has AVX-512 (NBH_ORACLE_ISA=v3, or
    use("v3"), selects the AVX2 build); both give the same results."""
    isa = _forced or os.environ.get("NBH_ORACLE_ISA", "")
    if isa != "v3" and _has_avx512() and os.path.exists(_LIB_V4):
        return _LIB_V4
    return _LIB_V3


def isas() -> list:
    """The builds this host can run: "v3", and "v4" where it has AVX-512."""
    return ["v3"] + (["v4"] if _has_avx512() and os.path.exists(_LIB_V4) else [])


def use(isa: str) -> None:
    """Load the "v3" (AVX2) or "v4" (AVX-512) build for later calls."""
    global _lib, _forced
    _forced = isa
    _lib = None


def variant() -> str:
    """Which build lib() loaded, with its compiler flags."""
    lib()
    return _variant


def build(force: bool = False) -> str:
    """Compile the oracle libraries in place (make -C oracle)."""
    if force or not (os.path.exists(_LIB_V3) and os.path.exists(_LIB_V4)):
        subprocess.run(["make", "-s", "-B" if force else "all", "-C", _HERE], check=True)
    return _LIB_V3


_d = ctypes.c_double
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_vp = ctypes.c_void_p


def lib():
    global _lib, _variant
    if _lib is None:
        if not os.path.exists(_LIB_V3):
            build()
        path = _lib_path()
        L = ctypes.CDLL(path)
        _variant = ("gcc -O3 -march=x86-64-v4 -mprefer-vector-width=512 -ffp-contract=off (AVX-512: the stage "
                    "loops 8 doubles a vector)" if path == _LIB_V4 else
                    "gcc -O3 -march=x86-64-v3 -ffp-contract=off (AVX2: the stage loops 4 doubles a vector)")
        sig = {
            "or_prime": (_u64, [_int]),
            "or_gen": (_u64, [_int]),
            "or_pinv": (_d, [_int]),
            "or_vec_reduce": (None, [_vp, _vp, _i64, _int]),
            "or_vec_mulmod3": (None, [_vp, _vp, _vp, _i64, _int]),
            "or_vec_mulmod4": (None, [_vp, _vp, _vp, _vp, _i64, _int]),
            "or_vec_minbal": (None, [_vp, _vp, _i64, _int]),
            "or_twiddles": (None, [_vp, _u64, _i64, _int]),
            "or_plan": (_vp, [_i64, _int]),
            "or_plan_build": (_vp, [_i64, _int]),
            "or_plan_free": (None, [_vp]),
            "or_plan_nstages": (_int, [_vp]),
            "or_plan_N2": (_i64, [_vp]),
            "or_plan_ninv": (_d, [_vp]),
            "or_plan_ninvp": (_d, [_vp]),
            "or_plan_stage": (None, [_vp, _int, _vp, _vp, _vp, _vp]),
            "or_plan_twiddle_row": (None, [_vp, _int, _int, _vp]),
            "or_fwd": (None, [_vp, _vp]),
            "or_rev": (None, [_vp, _vp]),
            "or_pack": (None, [_vp, _i64, _vp, _i64, _int, _i64]),
            "or_pointwise": (None, [_vp, _vp, _i64, _int]),
            "or_unpack2": (None, [_vp, _i64, _vp, _vp, _i64, _i64, _int, _d, _d, _d, _d]),
            "or_addmul_1": (_u64, [_vp, _vp, _i64, _u64]),
            "or_magnitude_bits": (_i64, [_vp, _i64]),
            "or_ntt_len": (_i64, [_i64]),
            "or_ntt_params": (_int, [_i64, _i64, _vp, _vp]),
            "or_mul": (_int, [_vp, _vp, _i64, _vp, _i64]),
            "or_sqr": (_int, [_vp, _vp, _i64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _u64a(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# -- constants (src/fpntt.jl:50-79) ------------------------------------------

def prime(k: int) -> int:
    return int(lib().or_prime(k))


def gen(k: int) -> int:
    return int(lib().or_gen(k))


def pinv(k: int) -> float:
    return float(lib().or_pinv(k))


# -- field arithmetic (src/fpntt.jl:81-121) ----------------------------------

def reduce(x, k: int) -> np.ndarray:
    x = _f64(x)
    out = np.empty_like(x)
    lib().or_vec_reduce(_ptr(out), _ptr(x), x.size, k)
    return out


def mulmod3(x, y, k: int) -> np.ndarray:
    x, y = np.broadcast_arrays(_f64(x), _f64(y))
    x, y = _f64(x), _f64(y)
    out = np.empty_like(x)
    lib().or_vec_mulmod3(_ptr(out), _ptr(x), _ptr(y), x.size, k)
    return out


def mulmod4(x, w, wp, k: int) -> np.ndarray:
    x, w, wp = (_f64(v) for v in np.broadcast_arrays(_f64(x), _f64(w), _f64(wp)))
    out = np.empty_like(x)
    lib().or_vec_mulmod4(_ptr(out), _ptr(x), _ptr(w), _ptr(wp), x.size, k)
    return out


def minbal(x, k: int) -> np.ndarray:
    x = _f64(x)
    out = np.empty_like(x)
    lib().or_vec_minbal(_ptr(out), _ptr(x), x.size, k)
    return out


def twiddles(r: int, n: int, k: int) -> np.ndarray:
    """fp_twiddles (src/fpntt.jl:173-186)."""
    out = np.empty(n, dtype=np.float64)
    lib().or_twiddles(_ptr(out), r, n, k)
    return out


# -- plans and transforms (src/fpntt.jl:191-271, 703-733) --------------------

class Plan:
    """FpNttPlan for (N, prime k); cached for the process like fp_ntt_plan."""

    def __init__(self, N: int, k: int):
        h = lib().or_plan(N, k)
        if not h:
            raise ValueError(f"invalid transform length {N}")
        self._h = ctypes.c_void_p(h)
        self.N = N
        self.k = k
        self.N2 = int(lib().or_plan_N2(self._h))
        self.ninv = float(lib().or_plan_ninv(self._h))
        self.ninvp = float(lib().or_plan_ninvp(self._h))
        self.stages = []
        for i in range(lib().or_plan_nstages(self._h)):
            m, span, Q = _int(), _i64(), _i64()
            rot = np.zeros(8, dtype=np.float64)
            lib().or_plan_stage(self._h, i, ctypes.byref(m), ctypes.byref(span),
                                ctypes.byref(Q), _ptr(rot))
            self.stages.append((m.value, span.value, Q.value, rot.copy()))

    def twiddle_row(self, i: int, t: int) -> np.ndarray:
        Q = self.stages[i][2]
        out = np.empty(Q, dtype=np.float64)
        lib().or_plan_twiddle_row(self._h, i, t, _ptr(out))
        return out

    def fwd(self, x) -> np.ndarray:
        """fp_ntt_fwd! (src/fpntt.jl:715-721); returns a new array."""
        y = _f64(x).copy()
        assert y.size == self.N
        lib().or_fwd(_ptr(y), self._h)
        return y

    def rev(self, x) -> np.ndarray:
        """fp_ntt_rev! (src/fpntt.jl:727-733); returns a new array."""
        y = _f64(x).copy()
        assert y.size == self.N
        lib().or_rev(_ptr(y), self._h)
        return y


# -- coefficient layer (src/fpntt.jl:744-925) --------------------------------

def pack(limbs, b: int, nch: int, N: int) -> np.ndarray:
    """fp_ntt_pack! (src/fpntt.jl:744-768)."""
    a = _u64a(limbs)
    out = np.empty(N, dtype=np.float64)
    lib().or_pack(_ptr(out), N, _ptr(a), a.size, b, nch)
    return out


def pointwise(xa, xb, k: int) -> np.ndarray:
    """fp_ntt_pointwise! (src/fpntt.jl:770-785); returns a new array."""
    a = _f64(xa).copy()
    b = _f64(xb)
    lib().or_pointwise(_ptr(a), _ptr(b), a.size, k)
    return a


def unpack2(rn: int, x1, x2, nconv: int, b: int, n1: float, n1p: float,
            n2: float, n2p: float) -> np.ndarray:
    """fp_ntt_unpack2! (src/fpntt.jl:869-925)."""
    x1 = _f64(x1)
    x2 = _f64(x2)
    r = np.zeros(rn, dtype=np.uint64)
    lib().or_unpack2(_ptr(r), rn, _ptr(x1), _ptr(x2), x1.size, nconv, b, n1, n1p, n2, n2p)
    return r


def addmul_1(r, a, b: int):
    """addmul_1! (src/kernels/mul.jl:32-70): r += a*b in place, returns the carry."""
    return int(lib().or_addmul_1(_ptr(r), _ptr(_u64a(a)), r.size, b))


# -- sizing (src/fpntt.jl:933-977, src/kernels/addsub.jl:150) ----------------

def magnitude_bits(a) -> int:
    a = _u64a(a)
    return int(lib().or_magnitude_bits(_ptr(a), a.size))


def ntt_len(T: int) -> int:
    """ntt_len (src/fpntt.jl:936-951); raises ValueError where it throws ArgumentError."""
    n = int(lib().or_ntt_len(T))
    if n < 0:
        raise ValueError(f"fp NTT length {T} exceeds the 315*2^34 the primes support")
    return n


def ntt_params(bits_a: int, bits_b: int):
    """fp_ntt_params (src/fpntt.jl:964-977) -> (b, N)."""
    b = _int()
    N = _i64()
    rc = lib().or_ntt_params(bits_a, bits_b, ctypes.byref(b), ctypes.byref(N))
    if rc:
        raise ValueError("unsupported operand size")
    return b.value, N.value


# -- engine (src/fpntt.jl:982-1037) -------------------------------------------

def mul(a, b) -> np.ndarray:
    """mul_fpntt2!: product limbs (len(a)+len(b))."""
    a = _u64a(a)
    b = _u64a(b)
    r = np.empty(a.size + b.size, dtype=np.uint64)
    if lib().or_mul(_ptr(r), _ptr(a), a.size, _ptr(b), b.size):
        raise ValueError("unsupported operand size")
    return r


def sqr(a) -> np.ndarray:
    """sqr_fpntt2!: square limbs (2*len(a))."""
    a = _u64a(a)
    r = np.empty(2 * a.size, dtype=np.uint64)
    if lib().or_sqr(_ptr(r), _ptr(a), a.size):
        raise ValueError("unsupported operand size")
    return r


# -- low short products (src/mul.jl:166-378) ----------------------------------
#
# mullo!/sqrlo! restated over Python integers: the same clamping, the same
# full-product cases, the same Mulders split ((11k) >> 4) with the two cross
# terms taken mod beta^(k-m) and the carry out of the top discarded.  The
# sub-products (Karatsuba / basecase / NTT in the reference) are exact integer
# products here; the truncated basecases are the product mod beta^k, which is
# what they compute.

MULLO_BASECASE_THRESHOLD = 140   # src/mul.jl:181-184
SQRLO_BASECASE_THRESHOLD = 224
MULLO_FULL_THRESHOLD = 224
SQRLO_FULL_THRESHOLD = 448
MUL_KARATSUBA_THRESHOLD = 29     # src/mul.jl:12


def _limbs_int(a, n: int) -> int:
    v = 0
    for i in range(n - 1, -1, -1):
        v = (v << 64) | int(a[i])
    return v


def _int_limbs(v: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    for i in range(n):
        out[i] = v & 0xFFFFFFFFFFFFFFFF
        v >>= 64
    return out


def mullo_split(k: int) -> int:
    """src/mul.jl:185."""
    return (11 * k) >> 4


def _mullo_int(A, la: int, B, lb: int, k: int) -> int:
    """mullo! (src/mul.jl:305-354) on limb arrays A, B; returns the low k
    limbs as an integer."""
    la, lb = min(la, k), min(lb, k)
    if la < lb:
        A, la, B, lb = B, lb, A, la
    beta_k = 1 << (64 * k)
    if la + lb <= k + 2:                               # full product in place
        return (_limbs_int(A, la) * _limbs_int(B, lb)) % beta_k
    if k < MULLO_BASECASE_THRESHOLD or lb < MUL_KARATSUBA_THRESHOLD or k >= MULLO_FULL_THRESHOLD:
        # truncated basecase, or full product with the low k limbs copied
        return (_limbs_int(A, la) * _limbs_int(B, lb)) % beta_k
    m = mullo_split(k)
    kk = k - m
    lla, llb = min(la, m), min(lb, m)
    r = (_limbs_int(A, lla) * _limbs_int(B, llb)) % beta_k
    if la > m:   # a_hi x b_lo, columns m..k
        c = _mullo_int(A[m:], la - m, B, min(lb, kk), kk)
        r = (r + (c << (64 * m))) % beta_k
    if lb > m:
        c = _mullo_int(B[m:], lb - m, A, min(la, kk), kk)
        r = (r + (c << (64 * m))) % beta_k
    return r


def mullo(a, b, k: int) -> np.ndarray:
    """mullo! (src/mul.jl:305-354): the k limbs of a*b mod beta^k."""
    a, b = _u64a(a), _u64a(b)
    return _int_limbs(_mullo_int(a, a.size, b, b.size, k), k)


def sqrlo(a, k: int) -> np.ndarray:
    """sqrlo! (src/mul.jl:360-378): the k limbs of a^2 mod beta^k."""
    a = _u64a(a)
    la = min(a.size, k)
    beta_k = 1 << (64 * k)
    A = _limbs_int(a, la)
    if 2 * la <= k + 2 or k < SQRLO_BASECASE_THRESHOLD or k >= SQRLO_FULL_THRESHOLD:
        return _int_limbs((A * A) % beta_k, k)
    m = mullo_split(k)
    kk = k - m
    lla = min(la, m)
    lo = _limbs_int(a, lla)
    r = (lo * lo) % beta_k
    if la > m:   # the cross term, added in twice; a_hi^2 starts at column 2m >= k
        c = _mullo_int(a[m:], la - m, a, min(la, kk), kk)
        r = (r + (c << (64 * m)) + (c << (64 * m))) % beta_k
    return _int_limbs(r, k)
