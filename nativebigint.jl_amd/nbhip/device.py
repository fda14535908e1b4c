"""Device-resident entry points (torch tensors as HBM buffers).

torch supplies device memory and the stream; all arithmetic runs in
libnbh.so's HIP kernels.  Limb buffers are int64 tensors holding the uint64
limbs bit for bit; transform buffers are float64 tensors.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from ._lib import (NBH_HGCD_ADJUST, NBH_HGCD_ADJUST_FULL, NBH_HGCD_FORWARD, check, init, lib, nbh_hgcd_job,
                   nbh_hgcd_level_job, nbh_job, nbh_spec_job)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)).to("cuda")


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint64).copy()


def empty_limbs(n: int) -> torch.Tensor:
    return torch.empty(n, dtype=torch.int64, device="cuda")


def _p(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def mul_dev(r: torch.Tensor, a: torch.Tensor, b: torch.Tensor, bits_a: int = -1, bits_b: int = -1) -> None:
    """r = a*b on the device, enqueued on torch's current stream."""
    init()
    assert r.numel() == a.numel() + b.numel()
    check(lib().nbh_mul_dev(_p(r), _p(a), a.numel(), _p(b), b.numel(), bits_a, bits_b, _stream()))


def sqr_dev(r: torch.Tensor, a: torch.Tensor, bits_a: int = -1) -> None:
    init()
    assert r.numel() == 2 * a.numel()
    check(lib().nbh_sqr_dev(_p(r), _p(a), a.numel(), bits_a, _stream()))


def mullo_dev(r: torch.Tensor, a: torch.Tensor, b: torch.Tensor, k: int, bits_a: int = -1,
              bits_b: int = -1) -> None:
    """r[0:k] = a*b mod beta^k (mullo!, src/mul.jl:305-354) on the device,
    enqueued on torch's current stream; r has at least k limbs."""
    init()
    assert r.numel() >= k
    check(lib().nbh_mullo_dev(_p(r), _p(a), a.numel(), _p(b), b.numel(), k, bits_a, bits_b, _stream()))


def sqrlo_dev(r: torch.Tensor, a: torch.Tensor, k: int, bits_a: int = -1) -> None:
    """r[0:k] = a^2 mod beta^k (sqrlo!, src/mul.jl:360-378)."""
    init()
    assert r.numel() >= k
    check(lib().nbh_sqrlo_dev(_p(r), _p(a), a.numel(), k, bits_a, _stream()))


class Batch:
    """A planned batch of independent products (device buffers).

    jobs: list of (r, a, b) tensors, b None for a square, optionally with
    (bits_a, bits_b) appended, and then k for a low short product r[0:k]
    (mullo!/sqrlo!).  run() enqueues every product on the current stream in a
    fixed number of launches."""

    def __init__(self, jobs):
        init()
        arr = (nbh_job * len(jobs))()
        self._keep = []
        for i, j in enumerate(jobs):
            r, a, b = j[0], j[1], j[2]
            bits_a, bits_b = (j[3], j[4]) if len(j) > 3 else (-1, -1)
            arr[i].k = j[5] if len(j) > 5 else 0
            arr[i].r = r.data_ptr()
            arr[i].a = a.data_ptr()
            arr[i].b = b.data_ptr() if b is not None else None
            arr[i].m = a.numel()
            arr[i].n = b.numel() if b is not None else a.numel()
            arr[i].bits_a = bits_a
            arr[i].bits_b = bits_b
            self._keep.append((r, a, b))
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        check(lib().nbh_batch_create(arr, len(jobs), ctypes.byref(h)))
        self._h = h

    def run(self, stream=None) -> None:
        s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
        check(lib().nbh_batch_run(self._h, s))

    def capture(self) -> None:
        """Capture the launch sequence into a HIP graph (replay() runs it)."""
        check(lib().nbh_batch_capture(self._h))

    def replay(self, stream=None) -> None:
        s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
        check(lib().nbh_batch_replay(self._h, s))

    def close(self) -> None:
        if self._h:
            lib().nbh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HgcdStep:
    """Batched 2x2 cofactor step (device tensors).

    instances: list of dicts with tensors m00, m01, m10, m11 (mn limbs each),
    u, v (n limbs each).  mode "forward": r0 = m00 u + m01 v, r1 = m10 u + m11 v
    (n+mn+1 limbs); mode "adjust": r0 = |m11 u - m01 v|, r1 = |m00 v - m10 u|
    (n+mn limbs) with signs s0, s1; mode "adjust_full": hgcd_matrix_adjust!
    (src/gcd.jl:398-437) with u, v the low p = n limbs and ra, rb (nn limbs
    each, also in the dict) the reduced tops: r0 = ra*2^(64p) + m11 u - m01 v,
    r1 = rb*2^(64p) + m00 v - m10 u, p + max(mn+1, nn) + 1 limbs, signs -1
    where the reference's nonnegativity assertion would fail.  Outputs are
    allocated here."""

    MODES = {"forward": NBH_HGCD_FORWARD, "adjust": NBH_HGCD_ADJUST, "adjust_full": NBH_HGCD_ADJUST_FULL}

    def __init__(self, instances, mode: str = "forward"):
        init()
        self.mode = self.MODES[mode]
        arr = (nbh_hgcd_job * len(instances))()
        self.out = []
        for i, q in enumerate(instances):
            mn, n = q["m00"].numel(), q["u"].numel()
            if self.mode == NBH_HGCD_ADJUST_FULL:
                nn = q["ra"].numel()
                assert q["rb"].numel() == nn
                lr = n + max(mn + 1, nn) + 1
                arr[i].ra, arr[i].rb, arr[i].nn = q["ra"].data_ptr(), q["rb"].data_ptr(), nn
            else:
                lr = n + mn + (1 if self.mode == NBH_HGCD_FORWARD else 0)
            r0 = torch.empty(lr, dtype=torch.int64, device="cuda")
            r1 = torch.empty(lr, dtype=torch.int64, device="cuda")
            sg = torch.zeros(2, dtype=torch.int32, device="cuda")
            self.out.append((r0, r1, sg, q))
            for k in ("m00", "m01", "m10", "m11", "u", "v"):
                setattr(arr[i], k, q[k].data_ptr())
            arr[i].mn, arr[i].n = mn, n
            arr[i].r0, arr[i].r1 = r0.data_ptr(), r1.data_ptr()
            arr[i].s0, arr[i].s1 = sg.data_ptr(), sg.data_ptr() + 4
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        check(lib().nbh_hgcd_create(arr, len(instances), self.mode, ctypes.byref(h)))
        self._h = h

    def run(self, stream=None) -> None:
        s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
        check(lib().nbh_hgcd_run(self._h, s))

    def close(self) -> None:
        if self._h:
            lib().nbh_hgcd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HgcdMatrixMul:
    """hgcd_matrix_mul! (src/gcd.jl:372-391), M <- M * M1, batched over many
    matrix pairs: row (x, y) of M becomes (x m1_00 + y m1_10, x m1_01 + y m1_11),
    i.e. one forward cofactor step with matrix M1 transposed and vector (x, y)
    (hgcd_mul2!, src/gcd.jl:360-367, twice).  pairs: list of (M, M1), dicts
    of device tensors m00..m11 (n and n1 limbs, zero-padded).  After run(),
    result(i) is the new M of pair i, entries of n + n1 + 1 limbs."""

    def __init__(self, pairs):
        inst = []
        for M, M1 in pairs:
            t = {"m00": M1["m00"], "m01": M1["m10"], "m10": M1["m01"], "m11": M1["m11"]}
            inst.append({**t, "u": M["m00"], "v": M["m01"]})
            inst.append({**t, "u": M["m10"], "v": M["m11"]})
        self.step = HgcdStep(inst, "forward")

    def run(self, stream=None) -> None:
        self.step.run(stream)

    def result(self, i: int) -> dict:
        (a0, a1, _, _), (b0, b1, _, _) = self.step.out[2 * i], self.step.out[2 * i + 1]
        return {"m00": a0, "m01": a1, "m10": b0, "m11": b1}

    def close(self) -> None:
        self.step.close()


class HgcdLevel:
    """One level of hgcd! (src/gcd.jl:546-554) for many instances, every
    buffer resident and updated in place: hgcd_matrix_adjust!(M1, a, b, p,
    ra, rb, nn) (src/gcd.jl:398-437) then hgcd_matrix_mul!(M, M1)
    (src/gcd.jl:372-391), run as one planned batch of 12 products per
    instance and one combine launch.

    instances: dicts of device tensors M = m00..m11 (capacity >= mn + m1n + 1
    limbs, the first mn used), M1 = n00..n11 (m1n limbs), a, b (capacity >=
    p + max(m1n + 1, nn) + 1, the low p limbs used), ra, rb (nn limbs), and
    ints mn, p.  After run(): a, b and the M entries hold the new values
    (zero-padded), signs[i] the adjust's signs (-1 where the reference's
    assertion would fail) and lens[i] = (normlen of a/b, new M.n)."""

    def __init__(self, instances):
        init()
        arr = (nbh_hgcd_level_job * len(instances))()
        count = len(instances)
        self.signs = torch.zeros((count, 2), dtype=torch.int32, device="cuda")
        self.lens = torch.zeros((count, 2), dtype=torch.int64, device="cuda")
        self._keep = instances
        for i, q in enumerate(instances):
            mn, p = int(q["mn"]), int(q["p"])
            m1n, nn = q["n00"].numel(), q["ra"].numel()
            L = p + max(m1n + 1, nn) + 1
            assert q["a"].numel() >= L and q["b"].numel() >= L and q["rb"].numel() == nn
            for k in ("m00", "m01", "m10", "m11"):
                assert q[k].numel() >= mn + m1n + 1
            for k in ("m00", "m01", "m10", "m11", "n00", "n01", "n10", "n11", "a", "b", "ra", "rb"):
                setattr(arr[i], k, q[k].data_ptr())
            arr[i].mn, arr[i].m1n, arr[i].p, arr[i].nn = mn, m1n, p, nn
            sp, lp = self.signs.data_ptr() + 8 * i, self.lens.data_ptr() + 16 * i
            arr[i].sa, arr[i].sb, arr[i].n_ab, arr[i].n_m = sp, sp + 4, lp, lp + 8
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        check(lib().nbh_hgcd_level_create(arr, count, ctypes.byref(h)))
        self._h = h

    def run(self, stream=None) -> None:
        s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else _stream()
        check(lib().nbh_hgcd_run(self._h, s))

    def close(self) -> None:
        if self._h:
            lib().nbh_hgcd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Spectrum:
    """Cached forward spectra of a fixed operand b (device tensor), for
    products with any a of up to bits_a_max bits: the modulus and reciprocal
    that barrett_reduce! (src/barrett.jl:44-81) multiplies by on every call.
    mul(r, a) runs a's transforms and the inverse only."""

    def __init__(self, b: torch.Tensor, bits_a_max: int, bits_b: int = -1):
        init()
        self.b = b
        h = ctypes.c_void_p()
        check(lib().nbh_spectrum_create(_p(b), b.numel(), bits_b, bits_a_max, _stream(), ctypes.byref(h)))
        self._h = h
        N, bch, n, amax = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().nbh_spectrum_info(h, ctypes.byref(N), ctypes.byref(bch), ctypes.byref(n), ctypes.byref(amax)))
        self.N, self.chunk_bits, self.n, self.bits_a_max = N.value, bch.value, n.value, amax.value

    @property
    def handle(self):
        return self._h

    def mul(self, r: torch.Tensor, a: torch.Tensor, bits_a: int = -1) -> None:
        """r = a*b on the device, enqueued on torch's current stream."""
        assert r.numel() == a.numel() + self.n
        check(lib().nbh_mul_spec_dev(_p(r), _p(a), a.numel(), bits_a, self._h, _stream()))

    def close(self) -> None:
        if self._h:
            lib().nbh_spectrum_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpecBatch(Batch):
    """A planned batch of products against cached spectra: jobs are
    (r, a, spectrum) or (r, a, spectrum, bits_a)."""

    def __init__(self, jobs):
        init()
        arr = (nbh_spec_job * len(jobs))()
        self._keep = []
        for i, j in enumerate(jobs):
            r, a, sp = j[0], j[1], j[2]
            arr[i].r = r.data_ptr()
            arr[i].a = a.data_ptr()
            arr[i].m = a.numel()
            arr[i].bits_a = j[3] if len(j) > 3 else -1
            arr[i].s = sp.handle
            self._keep.append((r, a, sp))
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        check(lib().nbh_spec_batch_create(arr, len(jobs), ctypes.byref(h)))
        self._h = h


# -- stage-level entry points (src/fpntt.jl:715, 727, 744, 770, 869) ---------

def fp_ntt_fwd(x: torch.Tensor, prime: int) -> None:
    init()
    check(lib().nbh_ntt_fwd_dev(_p(x), x.numel(), prime, _stream()))


def fp_ntt_rev(x: torch.Tensor, prime: int) -> None:
    init()
    check(lib().nbh_ntt_rev_dev(_p(x), x.numel(), prime, _stream()))


def fp_ntt_pack(x: torch.Tensor, limbs: torch.Tensor, b: int, nch: int) -> None:
    init()
    check(lib().nbh_pack_dev(_p(x), x.numel(), _p(limbs), limbs.numel(), b, nch, _stream()))


def fp_ntt_pointwise(xa: torch.Tensor, xb: torch.Tensor, prime: int) -> None:
    init()
    check(lib().nbh_pointwise_dev(_p(xa), _p(xb), xa.numel(), prime, _stream()))


def fp_ntt_unpack2(r: torch.Tensor, x1: torch.Tensor, x2: torch.Tensor, nconv: int, b: int) -> None:
    init()
    check(lib().nbh_unpack_dev(_p(r), r.numel(), _p(x1), _p(x2), x1.numel(), nconv, b, _stream()))


# -- plan introspection and tuning --------------------------------------------

def plan_stages(N: int, prime: int):
    init()
    nst = ctypes.c_int32()
    check(lib().nbh_plan_nstages(N, prime, ctypes.byref(nst)))
    out = []
    for i in range(nst.value):
        m, span, Q = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        rot = np.zeros(8, dtype=np.float64)
        check(lib().nbh_plan_stage(N, prime, i, ctypes.byref(m), ctypes.byref(span), ctypes.byref(Q),
                                   rot.ctypes.data_as(ctypes.c_void_p)))
        out.append((m.value, span.value, Q.value, rot))
    return out


def plan_twiddle_row(N: int, prime: int, i: int, t: int, Q: int) -> np.ndarray:
    out = np.empty(Q, dtype=np.float64)
    check(lib().nbh_plan_twiddle_row(N, prime, i, t, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def plan_ninv(N: int, prime: int):
    init()
    a, b = ctypes.c_double(), ctypes.c_double()
    check(lib().nbh_plan_ninv(N, prime, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def plan_groups(N: int):
    """[(La, Lb, S, C), ...] of the current schedule for N."""
    init()
    ng = ctypes.c_int32()
    buf = np.zeros(4 * 8, dtype=np.int64)
    check(lib().nbh_plan_groups(N, ctypes.byref(ng), buf.ctypes.data_as(ctypes.c_void_p)))
    return [tuple(int(v) for v in buf[4 * i:4 * i + 4]) for i in range(ng.value)]


def set_tuning(grp_lds: int, mid_max: int, cmin: int, cmax: int) -> None:
    init()
    check(lib().nbh_set_tuning(grp_lds, mid_max, cmin, cmax))


def get_tuning():
    init()
    v = [ctypes.c_int32() for _ in range(4)]
    check(lib().nbh_get_tuning(*[ctypes.byref(x) for x in v]))
    return tuple(x.value for x in v)


def set_tuning_table(entries, single: bool = False) -> None:
    """entries: [(Nmax, (grp_lds, mid_max, cmin, cmax)), ...] with increasing Nmax;
    single=True sets the table of products run alone."""
    init()
    nmax = np.array([e[0] for e in entries], dtype=np.int64)
    knobs = np.array([list(e[1]) for e in entries], dtype=np.int32).reshape(-1)
    f = lib().nbh_set_tuning_table_single if single else lib().nbh_set_tuning_table
    check(f(nmax.ctypes.data_as(ctypes.c_void_p) if len(entries) else None,
            knobs.ctypes.data_as(ctypes.c_void_p) if len(entries) else None, len(entries)))


def tuning_file_entries(single: bool = False):
    """The table the library loaded at init (tuning_<arch>.txt beside
    libnbh.so): the batch entries, or with single=True those of products run
    alone (the lines that start with "single")."""
    import os
    from ._lib import PKG_DIR
    path = os.path.join(PKG_DIR, f"tuning_{device_info()['arch'].split(':')[0]}.txt")
    if not os.path.exists(path):
        return []
    out = []
    for line in open(path):
        if not line.strip() or line.startswith("#"):
            continue
        w = line.split()
        if (w[0] == "single") != single:
            continue
        v = [int(x) for x in w[1:] if single] if single else [int(x) for x in w]
        out.append((v[0], tuple(v[1:5])))
    return out


def get_tuning_for_single(N: int):
    init()
    out = np.zeros(4, dtype=np.int32)
    check(lib().nbh_get_tuning_for_single(N, out.ctypes.data_as(ctypes.c_void_p)))
    return tuple(int(v) for v in out)


def get_tuning_for(N: int):
    init()
    out = np.zeros(4, dtype=np.int32)
    check(lib().nbh_get_tuning_for(N, out.ctypes.data_as(ctypes.c_void_p)))
    return tuple(int(v) for v in out)


def device_info():
    init()
    cus, lds, hbm = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
    name = ctypes.create_string_buffer(64)
    check(lib().nbh_device_info(ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(hbm), name, 64))
    return {"cus": cus.value, "lds_per_block": lds.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}


PROF_KINDS = ("group_fwd", "mid", "group_inv", "unpack", "carry_fix", "combine", "basecase", "small")


def set_streams(nside: int) -> None:
    """Side streams a batch step forks onto (0: launches in order)."""
    init()
    check(lib().nbh_set_streams(nside))


def profile_enable(on: bool = True) -> None:
    init()
    check(lib().nbh_profile_enable(1 if on else 0))


def profile_read() -> dict:
    """{kind: (total ms, launches)} since the last read (synchronises)."""
    init()
    ms = np.zeros(len(PROF_KINDS), dtype=np.float64)
    cnt = np.zeros(len(PROF_KINDS), dtype=np.int64)
    check(lib().nbh_profile_read(ms.ctypes.data_as(ctypes.c_void_p), cnt.ctypes.data_as(ctypes.c_void_p),
                                 len(PROF_KINDS)))
    return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROF_KINDS)}


def synchronize() -> None:
    init()
    check(lib().nbh_synchronize(None))
