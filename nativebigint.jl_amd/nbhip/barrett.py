"""Barrett reduction and powermod on the device, with the modulus and its
reciprocal held as cached spectra.

Mirrors barrett_mu! / barrett_reduce! (src/barrett.jl:24-81) and the
Barrett branch of powermod_limbs (src/algorithms.jl:17-107), batched over
instances that share the modulus: every product runs in libnbh.so's NTT
engine, and the products by mu and m skip those operands' transforms.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import device as D
from . import workloads as wl
from ._lib import check, init, lib, nbh_barrett_job


def barrett_mu(m: np.ndarray) -> tuple[np.ndarray, int]:
    """mu = floor(beta^(2k) / m) in k+2 limbs and its normalised length (k+1,
    or k+2 iff m = beta^(k-1)), as barrett_mu! returns (src/barrett.jl:24-31).
    One-off set-up per modulus, computed on the host."""
    k = int(m.size)
    M = wl.limbs_to_int(m)
    if M == 0 or int(m[-1]) == 0:
        raise ValueError("the modulus needs a nonzero top limb")
    mu = (1 << (128 * k)) // M
    lmu = max(1, (mu.bit_length() + 63) // 64)
    return wl.int_to_limbs(mu, k + 2), lmu


class Barrett:
    """Per-modulus context: m (k limbs, device) and mu's cached spectra."""

    def __init__(self, m: torch.Tensor):
        init()
        self.k = m.numel()
        mu, self.lmu = barrett_mu(D.to_host(m))
        self.m = m
        self.mu = D.to_dev(mu[:self.lmu])
        h = ctypes.c_void_p()
        check(lib().nbh_barrett_create(D._p(self.m), self.k, D._p(self.mu), self.lmu, D._stream(),
                                       ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            lib().nbh_barrett_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BarrettBatch:
    """Reductions r_i = t_i mod m for fixed buffers (t_i: 2k limbs, r_i: k
    limbs), planned once and re-run as the buffers' contents change."""

    def __init__(self, ctx: Barrett, pairs):
        init()
        arr = (nbh_barrett_job * len(pairs))()
        for i, (t, r) in enumerate(pairs):
            assert t.numel() == 2 * ctx.k and r.numel() == ctx.k
            arr[i].t = t.data_ptr()
            arr[i].r = r.data_ptr()
        self._keep = (ctx, list(pairs))
        h = ctypes.c_void_p()
        torch.cuda.synchronize()
        check(lib().nbh_barrett_batch_create(ctx.handle, arr, len(pairs), ctypes.byref(h)))
        self._h = h

    def run(self, stream=None) -> None:
        s = ctypes.c_void_p(stream.cuda_stream) if stream is not None else D._stream()
        check(lib().nbh_barrett_batch_run(self._h, s))

    def close(self) -> None:
        if self._h:
            lib().nbh_barrett_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def window_bits(nbits: int) -> int:
    """powermod_limbs' sliding-window width (src/algorithms.jl:27)."""
    return 1 if nbits <= 8 else 2 if nbits <= 24 else 3 if nbits <= 80 else 4 if nbits <= 240 else 5


def exponent_schedule(e: int):
    """The sequence of (op, table slot) powermod_limbs performs for e
    (src/algorithms.jl:78-100): ("load", j) starts the accumulator at table
    slot j, ("sqr", -1) squares it, ("mul", j) multiplies by slot j."""
    if e <= 0:
        raise ValueError("exponent must be positive")
    nbits = e.bit_length()
    w = window_bits(nbits)
    ops = []
    first = True
    i = nbits - 1
    while i >= 0:
        if not (e >> i) & 1:
            if not first:
                ops.append(("sqr", -1))
            i -= 1
        else:
            l = min(w, i + 1)
            while not (e >> (i - l + 1)) & 1:
                l -= 1
            val = 0
            for j in range(l):
                val = (val << 1) | ((e >> (i - j)) & 1)
            if first:
                ops.append(("load", val >> 1))
                first = False
            else:
                ops.extend([("sqr", -1)] * l)
                ops.append(("mul", val >> 1))
            i -= l
    return ops


class PowermodBatch:
    """b_i^e mod m for many bases b_i (0 < b_i < m, at most k limbs) and one
    exponent and modulus, every modular product a device product batch plus a
    Barrett reduction batch against the cached spectra of mu and m."""

    def __init__(self, m: np.ndarray, bases, e: int):
        init()
        self.k = k = int(m.size)
        self.e = e
        self.count = len(bases)
        self.ctx = Barrett(D.to_dev(m))
        self.w = window_bits(e.bit_length())
        tsize = 1 << (self.w - 1)
        self.ops = exponent_schedule(e)
        n = self.count
        full = 64 * k
        self.acc = [D.empty_limbs(k) for _ in range(n)]
        self.prod = [D.empty_limbs(2 * k) for _ in range(n)]
        self.bsq = [D.empty_limbs(k) for _ in range(n)]
        self.table = [[D.empty_limbs(k) for _ in range(n)] for _ in range(tsize)]
        for i, b in enumerate(bases):
            pad = np.zeros(k, dtype=np.uint64)
            pad[:b.size] = b
            self.table[0][i].copy_(D.to_dev(pad))
        # operands at full length (their contents change between runs)
        self.sq_acc = D.Batch([(self.prod[i], self.acc[i], None, full, full) for i in range(n)])
        self.mul_tab = [D.Batch([(self.prod[i], self.acc[i], self.table[j][i], full, full) for i in range(n)])
                        for j in range(tsize)]
        self.red_acc = BarrettBatch(self.ctx, list(zip(self.prod, self.acc)))
        self._setup = []
        if tsize > 1:
            self._setup.append(D.Batch([(self.prod[i], self.table[0][i], None, full, full) for i in range(n)]))
            self._setup.append(BarrettBatch(self.ctx, list(zip(self.prod, self.bsq))))
            for j in range(1, tsize):
                self._setup.append(D.Batch([(self.prod[i], self.table[j - 1][i], self.bsq[i], full, full)
                                            for i in range(n)]))
                self._setup.append(BarrettBatch(self.ctx, list(zip(self.prod, self.table[j]))))

    def modmuls(self) -> int:
        """Modular products per run (table set-up included)."""
        return len(self._setup) // 2 + sum(op != "load" for op, _ in self.ops)

    def run(self, stream=None) -> list:
        """Enqueue the whole exponentiation on `stream` (default: torch's
        current stream); returns the k-limb results (device tensors, valid
        once the stream has reached the end)."""
        stream = stream if stream is not None else torch.cuda.current_stream()
        with torch.cuda.stream(stream):
            for step in self._setup:
                step.run(stream)
            for op, j in self.ops:
                if op == "load":
                    for i in range(self.count):
                        self.acc[i].copy_(self.table[j][i])
                elif op == "sqr":
                    self.sq_acc.run(stream)
                    self.red_acc.run(stream)
                else:
                    self.mul_tab[j].run(stream)
                    self.red_acc.run(stream)
        return self.acc

    def capture(self) -> None:
        """Record one whole run (table set-up, every square, multiply and
        reduction, the accumulator loads) as a single HIP graph: a run is
        hundreds of short launches, and replay() submits them at once.  The
        graph holds kernels, events and device copies only."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        self.run(s)                     # eager once: every plan and table exists before capture
        s.synchronize()
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph, stream=s):
            self.run(s)
        torch.cuda.current_stream().wait_stream(s)

    def replay(self) -> list:
        """The captured run on torch's current stream (capture() first)."""
        self._graph.replay()
        return self.acc

    def close(self) -> None:
        self._graph = None
        for b in [self.sq_acc, self.red_acc, *self.mul_tab, *self._setup]:
            b.close()
        self.ctx.close()
