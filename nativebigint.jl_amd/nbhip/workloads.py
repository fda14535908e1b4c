"""Seeded synthetic operands and job lists for the multiplication workloads.

Limbs are i.i.d. uniform UInt64 drawn from a SplitMix64 stream per
(workload, job, operand), seed 0x4e42; a zero top limb is replaced by 1, so
every operand has its full limb length (the reference benches draw
`rand(UInt64, n)` the same way, bench/bench_kernels.jl:100-136).

Workloads (sizes in 64-bit limbs):
  single   4096 x 4096 product and a 4096-limb square
  large    1,048,576 x 1,048,576, 1,048,576 x 262,144, square of 4,194,304
  batch    4096 independent pairs, larger size log-uniform in [152, 65536],
           25% squarings, otherwise size ratio uniform in [1, 4] (smaller
           size at least 152, so the reference routes each to its NTT engine,
           src/mul.jl:388)
  hgcd     1024 instances of a 2x2 matrix of n/2-limb cofactors times a
           2-vector of n-limb operands, n log-uniform in [600, 32768]
  hgcdlevel 1024 instances of one hgcd! level's second half (the adjust by
           M1 and M <- M * M1), n0 log-uniform in [600, 32768]
  spectra  a fixed k-limb modulus m and (k+1)-limb reciprocal mu, and per
           Barrett reduction a (k+1)-limb q1 and q3 (barrett_reduce!,
           src/barrett.jl:44-81, multiplies q1*mu and q3*m)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SEED = 0x4E42
GAMMA = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB
MASK64 = (1 << 64) - 1

WORKLOAD_IDS = {"single": 1, "large": 2, "batch": 3, "hgcd": 4, "serve": 5, "spectra": 6, "powermod": 7, "small": 8,
                "test": 9, "shortprod": 10, "hgcdlevel": 11}


def _mix(z: int) -> int:
    z &= MASK64
    z = ((z ^ (z >> 30)) * _M1) & MASK64
    z = ((z ^ (z >> 27)) * _M2) & MASK64
    return z ^ (z >> 31)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First n outputs of SplitMix64 started at `seed` (vectorised, exact)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(_M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(_M2)
        z = z ^ (z >> np.uint64(31))
    return z


def stream_seed(workload: str | int, job: int, operand: int) -> int:
    wid = WORKLOAD_IDS[workload] if isinstance(workload, str) else int(workload)
    return _mix(SEED ^ (wid << 56) ^ (job << 8) ^ operand)


def operand(workload: str | int, job: int, which: int, n: int) -> np.ndarray:
    """n uniform limbs with a nonzero top limb."""
    a = splitmix64(stream_seed(workload, job, which), n)
    if n and a[-1] == 0:
        a[-1] = 1
    return a


@dataclass
class Job:
    m: int          # limbs of a (m >= n)
    n: int          # limbs of b (n == m for a square)
    sqr: bool


def _uniform01(seed: int, n: int) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def batch_jobs(count: int = 4096, lo: int = 152, hi: int = 65536, sqr_frac: float = 0.25,
               max_ratio: float = 4.0, workload: str = "batch") -> list[Job]:
    u = _uniform01(stream_seed(workload, 0xFFFFFF, 1), 3 * count).reshape(count, 3)
    jobs = []
    for k in range(count):
        L = int(math.floor(math.exp(math.log(lo) + u[k, 0] * (math.log(hi) - math.log(lo)))))
        L = max(lo, min(hi, L))
        if u[k, 1] < sqr_frac:
            jobs.append(Job(L, L, True))
        else:
            rho = 1.0 + u[k, 2] * (max_ratio - 1.0)
            jobs.append(Job(L, max(lo, int(L // rho)), False))
    return jobs


def serve_jobs(count: int, lo: int = 152, hi: int = 262144, sqr_frac: float = 0.3,
               unbal_frac: float = 0.2, max_ratio: float = 8.0, workload: str = "serve") -> list[Job]:
    """Serving stream: sizes log-uniform in [lo, hi]; 30% squares, 20%
    unbalanced with ratio uniform in [1, max_ratio], the rest balanced."""
    u = _uniform01(stream_seed(workload, 0xFFFFFE, 2), 3 * count).reshape(count, 3)
    jobs = []
    for k in range(count):
        L = int(math.floor(math.exp(math.log(lo) + u[k, 0] * (math.log(hi) - math.log(lo)))))
        L = max(lo, min(hi, L))
        if u[k, 1] < sqr_frac:
            jobs.append(Job(L, L, True))
        elif u[k, 1] < sqr_frac + unbal_frac:
            rho = 1.0 + u[k, 2] * (max_ratio - 1.0)
            jobs.append(Job(L, max(lo, int(L // rho)), False))
        else:
            jobs.append(Job(L, L, False))
    return jobs


def batch_operands(jobs: list[Job], workload: str = "batch"):
    """(a, b) limb arrays per job; b is None for squarings."""
    out = []
    for k, j in enumerate(jobs):
        a = operand(workload, k, 0, j.m)
        b = None if j.sqr else operand(workload, k, 1, j.n)
        out.append((a, b))
    return out


def hgcd_sizes(count: int = 1024, lo: int = 600, hi: int = 32768) -> list[int]:
    u = _uniform01(stream_seed("hgcd", 0xFFFFFF, 1), count)
    out = []
    for k in range(count):
        n = int(math.floor(math.exp(math.log(lo) + u[k] * (math.log(hi) - math.log(lo)))))
        out.append(max(lo, min(hi, n)))
    return out


def hgcd_instance(k: int, n: int, workload: str = "hgcd") -> dict:
    """Instance k of the cofactor-step workload: four n/2-limb cofactors and
    two n-limb operands, uniform limbs with nonzero top limbs."""
    mn = max(1, n // 2)
    names = ("m00", "m01", "m10", "m11", "u", "v")
    return {name: operand(workload, k, i, mn if i < 4 else n) for i, name in enumerate(names)}


def hgcd_level_shape(n0: int) -> dict:
    """Buffer sizes of the second half of one hgcd! level on n0-limb inputs
    (src/gcd.jl:525-554): after the first recursion and the stray steps the
    values have n = 3 n0 / 4 limbs, s = n0 / 2 + 1, the second recursion runs
    on the top n - p limbs with p = 2 s - n + 1 and returns M1 (about (n - p)/2
    limbs) and the reduced tops ra, rb (nn = (n - p)/2 + 1 limbs); M, from the
    first recursion on the top n0 / 2 limbs, has about n0 / 4."""
    s = n0 // 2 + 1
    n = 3 * n0 // 4
    p = 2 * s - n + 1
    m1n = max(1, (n - p) // 2)
    nn = (n - p) // 2 + 1
    mn = max(1, n0 // 4)
    return {"n": n, "p": p, "mn": mn, "m1n": m1n, "nn": nn, "L": p + max(m1n + 1, nn) + 1,
            "Lm": mn + m1n + 1}


def hgcd_level_instance(k: int, n0: int, workload: str = "hgcdlevel") -> dict:
    """Instance k of the level workload: M (mn-limb entries), M1 (m1n-limb
    entries), a, b (n limbs), ra, rb (nn limbs), uniform limbs with nonzero
    top limbs.  Entries are host arrays at their own lengths; the caller pads
    them to the capacities in hgcd_level_shape."""
    sh = hgcd_level_shape(n0)
    out = {}
    names = [("m00", "mn"), ("m01", "mn"), ("m10", "mn"), ("m11", "mn"), ("n00", "m1n"), ("n01", "m1n"),
             ("n10", "m1n"), ("n11", "m1n"), ("a", "n"), ("b", "n"), ("ra", "nn"), ("rb", "nn")]
    for i, (name, size) in enumerate(names):
        out[name] = operand(workload, k, i, sh[size])
    return out


def barrett_operands(count: int = 1024, k: int = 4096, workload: str = "spectra"):
    """(m, mu, [q1], [q3]): m has k limbs, mu and every q1, q3 k+1 limbs."""
    m = operand(workload, 0xFFFFF0, 0, k)
    mu = operand(workload, 0xFFFFF0, 1, k + 1)
    q1 = [operand(workload, i, 0, k + 1) for i in range(count)]
    q3 = [operand(workload, i, 1, k + 1) for i in range(count)]
    return m, mu, q1, q3


def limbs_to_int(a) -> int:
    return int.from_bytes(np.ascontiguousarray(a, dtype="<u8").tobytes(), "little")


def int_to_limbs(x: int, n: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)
