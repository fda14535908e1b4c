"""Pins the CPU oracle (oracle/fpntt_oracle.c) before anything is checked against it.

Each check is independent of the oracle's own formulas: Python big-int
arithmetic, closed forms, brute-force DFT/convolution, the reference test's
literal values, libgmp.  The structure follows test/test_fpntt.jl.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import gmp_ref
from tests.helpers import ints, rand_big_style, wl

P1 = 65205 * 2**34 + 1
P2 = 4095 * 2**38 + 1
PR = (P1, P2)


def canon(x: float, p: int) -> int:
    return int(x) % p


# -- constants (test/test_fpntt.jl:20-25, 163-168) -----------------------------

def test_primes_and_generators():
    assert O.prime(0) == P1 and O.prime(1) == P2
    assert float(P1) == P1 and float(P2) == P2          # exact doubles
    assert (P1 - 1) % (315 << 34) == 0 and (P2 - 1) % (315 << 38) == 0
    assert P1 < 2**50 and P2 < 2**50
    # generator: smallest g in 3..63 of full order (brute force over factors)
    for k, (p, facs) in enumerate(((P1, (2, 3, 5, 7, 23)), (P2, (2, 3, 5, 7, 13)))):
        q = p - 1
        for f in facs:
            while q % f == 0:
                q //= f
        assert q == 1                                   # facs is complete
        g = next(g for g in range(3, 64) if all(pow(g, (p - 1) // f, p) != 1 for f in facs))
        assert O.gen(k) == g
    assert (O.gen(0), O.gen(1)) == (33, 11)
    assert O.pinv(0) == 1.0 / P1 and O.pinv(1) == 1.0 / P2
    assert pow(P1, -1, P2) == 208 and 208 * P1 == 207 * P2 + 1   # Garner constant
    for logN in (1, 5, 20, 34):                         # roots of exact order
        w = pow(33, (P1 - 1) >> logN, P1)
        assert pow(w, 1 << logN, P1) == 1 and pow(w, 1 << (logN - 1), P1) == P1 - 1


# -- field arithmetic over the lazy range (test/test_fpntt.jl:26-68) -----------

def _edge_sets(k, rng):
    p = float(PR[k])
    half = (p - 1) / 2
    ws = [0.0, 1.0, -1.0, 2.0, 2.0**33, half, -half]
    ws += [float(rng.choice([-1, 1]) * rng.integers(0, (PR[k] - 1) >> 1)) for _ in range(60)]
    xs = [0.0, 1.0, -1.0, p - 1, -(p - 1), p, -p, 2 * p, -2 * p, 4 * p - 4, -(4 * p - 4), 4 * p, -4 * p]
    xs += [float(rng.choice([-4, -2, -1, 1, 2, 4]) * rng.integers(0, PR[k] - 1)) for _ in range(200)]
    return np.array(ws), np.array(xs)


@pytest.mark.parametrize("k", [0, 1])
def test_field_ops_lazy_range(k):
    rng = np.random.default_rng(0xF94 + k)
    p = PR[k]
    ws, xs = _edge_sets(k, rng)
    X, W = np.meshgrid(xs, ws)
    X, W = X.ravel(), W.ravel()
    r = O.mulmod4(X, W, W / float(p), k)                # Shoup form, |x| <= 4p, |w| <= p/2
    assert np.all(np.abs(r) < p)
    for x, w, v in zip(X[::7], W[::7], r[::7]):
        assert int(v) % p == (int(x) * int(w)) % p
    exact = np.array([(int(x) * int(w)) % p for x, w in zip(X, W)], dtype=object)
    assert all(int(v) % p == e for v, e in zip(r, exact))
    r3 = O.mulmod3(X, W, k)
    assert np.all(np.abs(r3) < p)
    assert all(int(v) % p == e for v, e in zip(r3, exact))
    small = xs[np.abs(xs) <= p]                         # value x value, |x y| <= p^2
    A, B = np.meshgrid(small, small)
    rv = O.mulmod3(A.ravel(), B.ravel(), k)
    assert np.all(np.abs(rv) < p)
    assert all(int(v) % p == (int(a) * int(b)) % p for v, a, b in zip(rv, A.ravel(), B.ravel()))
    rr = O.reduce(2 * xs, k)                            # up to 8p
    assert np.all(np.abs(rr) < p)
    assert all(int(v) % p == (2 * int(x)) % p for v, x in zip(rr, xs))
    cand = np.concatenate([ws, [(p - 1) / 2 + 1, -((p - 1) / 2 + 1), p - 1, -(p - 1)]])
    mb = O.minbal(cand, k)
    assert np.all(np.abs(mb) <= p / 2)
    assert all(int(v) % p == int(x) % p for v, x in zip(mb, cand))


def test_round_is_exact_nearest():
    # fp_reduce of exact multiples and near-multiples of p lands in (-p, p)
    for k in (0, 1):
        p = PR[k]
        xs = np.array([float(q * p + d) for q in range(-2**20, 2**20, 2**15 + 17) for d in (-3, 0, 5)])
        r = O.reduce(xs, k)
        assert np.all(np.abs(r) < p)
        assert all(int(v) % p == int(x) % p for v, x in zip(r, xs))


# -- twiddles and plan constants (src/fpntt.jl:127-271), pinned by closed form -

def _minbal_int(v, p):
    v %= p
    return v - p if v > p // 2 else v


@pytest.mark.parametrize("k", [0, 1])
def test_twiddles_closed_form(k):
    p = PR[k]
    for N in (4, 64, 12, 1536, 960, 28, 36, 4032):
        r = pow(O.gen(k), (p - 1) // N, p)
        t = O.twiddles(r, N, k)
        want = [_minbal_int(pow(r, j, p), p) for j in range(N)]
        assert [int(v) for v in t] == want


def _stage_roots(N, k):
    """Root of unity each stage's twiddle rows are powers of, in stage order."""
    p = PR[k]
    w = pow(O.gen(k), (p - 1) // N, p)
    m = N >> (N & -N).bit_length() - 1
    out = []
    span, r = N, w
    for mf in (7, 5, 3, 3):
        if m % mf == 0:
            out.append((mf, span, r))
            r = pow(r, mf, p)
            span //= mf
            m //= mf
    L = span
    while L >= 4:
        out.append((4, L, pow(r, span // L, p)))
        L >>= 2
    return out, span, r


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", [4, 8, 16, 64, 12, 20, 48, 160, 240, 28, 84, 140, 420, 36, 144, 720, 1260, 4032, 1680])
def test_plan_constants_closed_form(k, N):
    p = PR[k]
    P = O.Plan(N, k)
    roots, N2, rtow = _stage_roots(N, k)
    assert P.N2 == N2
    assert len(P.stages) == len(roots)
    inv2, inv3, inv4 = pow(2, -1, p), pow(3, -1, p), pow(4, -1, p)
    for i, ((m, span, Q, rot), (mm, L, wL)) in enumerate(zip(P.stages, roots)):
        assert m == mm and Q == L // m
        # every twiddle row t is (w_L^t)^j, minimal balanced
        for t in range(m - 1):
            row = P.twiddle_row(i, t)
            assert [int(v) for v in row] == [_minbal_int(pow(wL, (t + 1) * j, p), p) for j in range(Q)]
        c = pow(wL, L // m, p)                          # primitive m-th root of the combine
        assert pow(c, m, p) == 1 and c != 1
        if m == 4:
            assert int(rot[0]) % p == pow(rtow, N2 // 4, p) and abs(rot[0]) < p / 2 + 1
        elif m == 3:
            assert int(rot[0]) % p == c and abs(rot[0]) < p / 2 + 1
            assert int(rot[1]) % p == c * c % p
        elif m == 5:
            want = [-inv4, (c + c**4 - c**2 - c**3) * inv4, (c - c**4) * inv2, (c**2 - c**3) * inv2]
            assert [int(v) for v in rot[:4]] == [_minbal_int(x, p) for x in want]
        else:
            cj = [pow(c, j, p) for j in range(7)]
            A = lambda j: (cj[j] + cj[7 - j]) * inv2
            B = lambda j: (cj[j] - cj[7 - j]) * inv2
            a0, a1, a2 = A(1), A(3), A(2)
            b0, b1, b2 = B(1), B(3), B(2)
            want = [(a0 + a1 + a2) * inv3, (a0 - a2) * inv3, (a1 - a2) * inv3,
                    (a0 - a2) * inv3 + (a1 - a2) * inv3, (b0 - b1 + b2) * inv3,
                    (b0 - b2) * inv3, (b1 + b2) * inv3, (b0 - b2) * inv3 + (b1 + b2) * inv3]
            assert [int(v) for v in rot] == [_minbal_int(x, p) for x in want]
    assert int(P.ninv) % p == pow(N, -1, p) and abs(P.ninv) < p / 2 + 1
    assert P.ninvp == P.ninv / p


# -- transform (test/test_fpntt.jl:84-116, 163-181) ---------------------------

def _digit_reversed_freq(N, k):
    """Output slot -> frequency of the DIF network: slot digits (t1, t2, ...),
    most significant first with radices in stage order, give frequency
    t1 + m1 (t2 + m2 (...))."""
    roots, N2, _ = _stage_roots(N, k)
    radices = [m for (m, _, _) in roots]
    if (N2.bit_length() - 1) % 2 == 1:
        radices.append(2)
    freq = np.zeros(N, dtype=np.int64)
    for slot in range(N):
        rem, span, f, mult = slot, N, 0, 1
        for m in radices:
            Q = span // m
            t, rem = divmod(rem, Q)
            f += t * mult
            mult *= m
            span = Q
        freq[slot] = f
    return freq


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", [4, 8, 16, 32, 64, 128, 12, 20, 48, 96, 160, 240, 28, 56, 84, 140, 36, 144, 180, 252])
def test_forward_is_digit_reversed_dft(k, N):
    p = PR[k]
    rng = np.random.default_rng(N * 7 + k)
    x = rng.integers(0, p, N)
    y = O.Plan(N, k).fwd(x.astype(np.float64))
    w = pow(O.gen(k), (p - 1) // N, p)
    freq = _digit_reversed_freq(N, k)
    xs = [int(v) for v in x]
    for slot in range(N):
        f = int(freq[slot])
        X = sum(xs[j] * pow(w, j * f, p) for j in range(N)) % p
        assert canon(y[slot], p) == X
    assert np.all(np.abs(y) <= 1.5 * p)


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", [4, 8, 64, 512, 1024, 12, 20, 48, 96, 160, 240, 1536, 960, 3840,
                               28, 56, 84, 140, 420, 1680, 36, 144, 720, 1260, 4032, 2**16, 7 * 2**12])
def test_round_trip(k, N):
    p = PR[k]
    rng = np.random.default_rng(N + 31 * k)
    x = rng.integers(0, p, N)
    P = O.Plan(N, k)
    y = P.fwd(x.astype(np.float64))
    assert not np.array_equal(y, x.astype(np.float64))
    z = P.rev(y)
    idx = (N - np.arange(N)) % N
    want = (np.array([int(v) for v in x], dtype=object)[idx] * N) % p
    got = np.array([int(v) % p for v in z], dtype=object)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("k", [0, 1])
@pytest.mark.parametrize("N", [4, 8, 12, 20, 28, 36, 48, 64, 96, 140, 160, 240])
def test_cyclic_convolution(k, N):
    p = PR[k]
    rng = np.random.default_rng(0xF18 + N)
    a = [int(v) for v in rng.integers(0, p, N)]
    b = [int(v) for v in rng.integers(0, p, N)]
    c = [0] * N
    for i in range(N):
        for j in range(N):
            c[(i + j) % N] += a[i] * b[j]
    P = O.Plan(N, k)
    fa = P.fwd(np.array(a, dtype=np.float64))
    fb = P.fwd(np.array(b, dtype=np.float64))
    fc = P.rev(O.pointwise(fa, fb, k))
    idx = (N - np.arange(N)) % N
    got = O.mulmod4(fc[idx], P.ninv, P.ninvp, k)
    assert [int(v) % p for v in got] == [v % p for v in c]


# -- sizing (test/test_fpntt.jl:205-255) --------------------------------------

def _want_len(T):
    ms = (1, 3, 5, 9) if T <= 1 << 14 else (1, 3, 5, 7, 9, 15, 21, 105) if T <= 1 << 20 else \
        (1, 3, 5, 7, 9, 15, 21, 35, 45, 63, 105, 315)
    return min(m << k for m in ms for k in range(2, 35) if (m << k) >= T)


def test_ntt_len_kats():
    Ts = [1, 4, 5, 100, 1000, 1023, 1025, 13000, 15359, 15360, 20000, 1 << 20,
          (15 << 14) - 100, 15 << 14, (15 << 14) + 1, (15 << 20) - 1, 15 << 20,
          (21 << 14) - 1, 21 << 14, (35 << 14) - 1, 35 << 14, (105 << 14) - 1, 105 << 14,
          (7 << 10) - 1, 7 << 10, (3 << 22) - 1, 3 << 22, (5 << 22) - 1, 5 << 22,
          (7 << 22) - 1, 7 << 22, 1 << 22, 1 << 25]
    for T in Ts:
        n = O.ntt_len(T)
        assert n == _want_len(T) and n >= T
        assert (P1 - 1) % n == 0 and (P2 - 1) % n == 0
    # literal values from the reference test
    assert O.ntt_len(15 << 14) == 15 << 14
    assert O.ntt_len(15 << 13) == 15 << 13
    assert O.ntt_len(7 << 10) == 1 << 13
    assert O.ntt_len(7 << 14) == 7 << 14
    assert O.ntt_len(9 << 10) == 9 << 10
    assert O.ntt_len((1 << 34) + 1) == 35 << 29
    assert O.ntt_len(1 << 34) == 1 << 34
    assert O.ntt_len(33 << 34) == 35 << 34
    assert O.ntt_len(53 << 34) == 63 << 34
    assert O.ntt_len(315 << 34) == 315 << 34
    with pytest.raises(ValueError):
        O.ntt_len((315 << 34) + 1)


def test_ntt_params_bound_and_kats():
    pp = P1 * P2 - 1
    for bits_a, bits_b in [(64, 64), (64 * 8, 64 * 8), (64 * 152, 64 * 152), (64 * 4096, 64 * 4096),
                           (64 * 2**20, 64 * 2**20), (64 * 2**20, 64 * 2**18), (64 * 2**22, 64 * 2**22),
                           (100, 1), (64 * 144000, 64 * 144000)]:
        b, N = O.ntt_params(bits_a, bits_b)
        nca, ncb = -(-bits_a // b), -(-bits_b // b)
        assert min(nca, ncb) * (2**b - 1) ** 2 <= pp           # exactness bound
        if b < 49:
            nca1, ncb1 = -(-bits_a // (b + 1)), -(-bits_b // (b + 1))
            assert min(nca1, ncb1) * (2**(b + 1) - 1) ** 2 > pp  # b is the largest
        assert N == _want_len(nca + ncb - 1)
    assert O.ntt_params(64 * 4096, 64 * 4096) == (43, 12288)
    assert O.ntt_params(64 * 2**20, 64 * 2**20) == (39, 7 * 2**19)
    assert O.ntt_params(64 * 2**20, 64 * 2**18) == (40, 2**21)
    assert O.ntt_params(64 * 2**22, 64 * 2**22) == (38, 7 * 2**21)
    b, N = O.ntt_params(64 * 144000, 64 * 144000)
    assert N == 7 * 2**16                                  # test/test_fpntt.jl:308-316


def test_magnitude_bits_and_addmul():
    a = np.array([5, 0, 1 << 40], dtype=np.uint64)
    assert O.magnitude_bits(a) == 128 + 41
    assert O.magnitude_bits(np.array([0, 0], dtype=np.uint64)) == 64   # upper bound on zero top
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 20):
        for _ in range(20):
            a = rng.integers(0, 2**64, n, dtype=np.uint64)
            r = rng.integers(0, 2**64, n, dtype=np.uint64)
            b = int(rng.integers(0, 2**64, dtype=np.uint64))
            r0 = ints(r)
            c = O.addmul_1(r, a, b)
            assert ints(r) + (c << (64 * n)) == r0 + ints(a) * b


# -- Garner unpack vs big-int CRT (test/test_fpntt.jl:257-286) -----------------

@pytest.mark.parametrize("b", [33, 40, 44, 48])
@pytest.mark.parametrize("nconv", [1, 7, 64, 200])
def test_unpack_vs_crt(b, nconv):
    rng = np.random.default_rng(0x6A3 + 1000 * b + nconv)
    cs = [int.from_bytes(rng.bytes(16), "little") % (P1 * P2) for _ in range(nconv)]
    N = nconv + int(rng.integers(0, 6))
    x1 = np.zeros(N)
    x2 = np.zeros(N)
    for i, c in enumerate(cs):
        idx = (N - i) % N
        x1[idx] = float((c * N) % P1) + int(rng.integers(-2, 3)) * float(P1)
        x2[idx] = float((c * N) % P2) + int(rng.integers(-2, 3)) * float(P2)
    n1 = float(O.reduce([float(pow(N, -1, P1))], 0)[0])
    n2 = float(O.reduce([float(pow(N, -1, P2))], 1)[0])
    rn = -(-(99 + b * (nconv - 1)) // 64) + 1
    r = O.unpack2(rn, x1, x2, nconv, b, n1, n1 / P1, n2, n2 / P2)
    assert ints(r) == sum(c << (b * i) for i, c in enumerate(cs))


# -- products (test/test_fpntt.jl:288-331, test/test_algorithms.jl:109-128) ---

@pytest.mark.parametrize("na,nb", [(1, 1), (5, 3), (16, 16), (100, 100), (255, 257), (500, 300),
                                   (601, 373), (1024, 1024), (2500, 2500), (5000, 4900),
                                   (151, 151), (152, 152), (153, 153), (456, 76), (381, 153)])
def test_mul_differential(na, nb):
    rng = np.random.default_rng(0x91C4 + na * 7919 + nb)
    for _ in range(2 if na * nb > 10**5 else 6):
        a = rand_big_style(rng, na)
        b = rand_big_style(rng, nb)
        assert ints(O.mul(a, b)) == ints(a) * ints(b)


def test_all_ones_reach_b49():
    for na in range(1, 9):
        for nb in range(1, na + 1):
            a = np.full(na, 2**64 - 1, dtype=np.uint64)
            b = np.full(nb, 2**64 - 1, dtype=np.uint64)
            assert ints(O.mul(a, b)) == ints(a) * ints(b)
    assert O.ntt_params(64, 64)[0] == 49


@pytest.mark.parametrize("n", [1, 7, 100, 159, 160, 161, 321, 1024, 3000])
def test_sqr_differential(n):
    rng = np.random.default_rng(0x3F17 + n)
    a = rand_big_style(rng, n)
    assert ints(O.sqr(a)) == ints(a) ** 2
    ones = np.full(6, 2**64 - 1, dtype=np.uint64)
    assert ints(O.sqr(ones)) == ints(ones) ** 2


def test_radix7_length_by_residues():
    """144000-limb product at N = 7*2^16, checked mod random 63-bit moduli
    (test/test_fpntt.jl:308-325)."""
    n = 144_000
    a = wl.operand("test", 1, 0, n)
    b = wl.operand("test", 1, 1, n)
    r = O.mul(a, b)
    A, B, R = ints(a), ints(b), ints(r)
    rng = np.random.default_rng(7)
    for m in rng.integers(2**62, 2**63 - 1, 4):
        m = int(m)
        assert R % m == (A % m) * (B % m) % m


@pytest.mark.skipif(not gmp_ref.available(), reason="libgmp not present")
def test_single_workload_vs_gmp():
    a = wl.operand("single", 0, 0, 4096)
    b = wl.operand("single", 0, 1, 4096)
    assert np.array_equal(O.mul(a, b), gmp_ref.mul(a, b))
    assert np.array_equal(O.sqr(a), gmp_ref.sqr(a))
