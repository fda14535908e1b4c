"""Continuous-batching multiplication server on one GPU.

Requests carry HBM-resident operands.  The serving loop (`Server.poll`)
  - retires micro-batches whose completion event has fired (their arena space
    is freed, their requests get a completion time),
  - admits queued requests, oldest first, while their results and workspace
    fit the free part of the HBM budget (one contiguous arena block per
    micro-batch; a request that can never fit is rejected with
    NBH_ERR_BUDGET),
  - plans the admitted requests as one micro-batch *inside* that block
    (nbh_batch_create_in: no device allocation, no host synchronisation) and
    launches it on a free stream while earlier micro-batches are still running.
So requests that arrive while others are on the GPU start as soon as a stream
and memory are free, and their work overlaps on the device.  Plans and
twiddle tables stay resident across batches (the library's caches).

Latency of a request = time its completion was observed by the loop minus its
arrival time (the device finished at or before the observation).
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from ._lib import NBH_ERR_BUDGET, NbhError, check, init, lib, nbh_job

ALIGN = 256


@dataclass
class Request:
    a: torch.Tensor                      # operand limbs (int64 view of uint64), device
    b: torch.Tensor | None               # None for a square
    bits_a: int
    bits_b: int
    t_arrival: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    off: int = -1                        # byte offset of the result in the arena while in flight
    keep: bool = False                   # copy the result out on completion
    out: torch.Tensor | None = None
    rid: int = 0

    def __post_init__(self):
        self.dev_a = self.a.data_ptr()
        self.dev_b = self.b.data_ptr() if self.b is not None else 0
        self.m = self.a.numel()
        self.n = self.m if self.b is None else self.b.numel()
        self.limbs = self.m + self.n
        # the nbh_job record (r filled in at admission) and the sizing key
        self.row = np.array([0, self.dev_a, self.dev_b, self.m, self.n, self.bits_a, self.bits_b, 0], dtype=np.int64)
        self.key = (self.m, None if self.b is None else self.n, self.bits_a, self.bits_b)


class Arena:
    """First-fit allocator of byte ranges in one device buffer."""

    def __init__(self, size: int):
        self.size = size
        self.free = [(0, size)]          # sorted (offset, length)

    def alloc(self, n: int):
        n = (n + ALIGN - 1) // ALIGN * ALIGN
        for i, (off, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + n, ln - n)
                return off, n
        return None

    def release(self, off: int, n: int):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged

    @property
    def largest_free(self) -> int:
        return max((ln for _, ln in self.free), default=0)


@dataclass
class _Flight:
    reqs: list
    handle: ctypes.c_void_p
    event: torch.cuda.Event
    block: tuple
    stream: int
    rec: np.ndarray   # the job records (kept until the batch is released)


@dataclass
class Server:
    budget_bytes: int
    nstreams: int = 4
    max_batch: int = 512
    pending: deque = field(default_factory=deque)
    done: list = field(default_factory=list)
    rejected: list = field(default_factory=list)

    def __post_init__(self):
        init()
        self.mem = torch.empty(self.budget_bytes, dtype=torch.uint8, device="cuda")
        self.base = self.mem.data_ptr()
        self.arena = Arena(self.budget_bytes)
        self.streams = [torch.cuda.Stream() for _ in range(self.nstreams)]
        self.busy = [None] * self.nstreams
        self.batches = 0
        self._rid = 0
        self._need_cache = {}
        self.host_admit_s = 0.0   # host time spent planning and launching micro-batches

    # -- requests ---------------------------------------------------------------
    def submit(self, req: Request, t: float | None = None) -> Request:
        req.t_arrival = time.perf_counter() if t is None else t
        req.rid = self._rid
        self._rid += 1
        self.pending.append(req)
        return req

    def idle(self) -> bool:
        return not self.pending and all(b is None for b in self.busy)

    # -- the loop -----------------------------------------------------------------
    def _jobs(self, reqs, results_off):
        """nbh_job records (r, a, b, m, n, bits_a, bits_b, k=0) as one int64 array,
        the results laid out back to back (ALIGN-aligned) from results_off;
        returns the records, the results' bytes and their offsets."""
        rec = np.stack([r.row for r in reqs])
        size = (8 * (rec[:, 3] + rec[:, 4]) + ALIGN - 1) // ALIGN * ALIGN
        offs = results_off + np.cumsum(size) - size
        rec[:, 0] = self.base + offs
        return rec, int(size.sum()), offs

    @staticmethod
    def _ptr(rec):
        return rec.ctypes.data_as(ctypes.POINTER(nbh_job))

    def _need1(self, r):
        """Arena bytes of one request alone (result + workspace + tables).  A
        micro-batch needs at most the sum over its requests (its tables are
        shared and aligned once), so admission sums these."""
        key = r.key
        v = self._need_cache.get(key)
        if v is None:
            rec, res_bytes, _ = self._jobs([r], 0)
            ws = ctypes.c_int64()
            check(lib().nbh_batch_arena_bytes(self._ptr(rec), 1, ctypes.byref(ws)))
            v = res_bytes + ws.value + ALIGN
            self._need_cache[key] = v
        return v

    def prepare(self, requests) -> None:
        """Size every distinct request shape once (building the plans and
        twiddle tables of their transform lengths), as a server does at
        start-up, so admission never builds a plan."""
        for r in requests:
            self._need1(r)

    def _retire(self):
        now = time.perf_counter()
        for k, f in enumerate(self.busy):
            if f is not None and f.event.query():
                for r in f.reqs:
                    r.t_done = now
                    if r.keep:
                        with torch.cuda.stream(self.streams[k]):
                            r.out = self.result_view(r).clone()
                    r.off = -1
                self.done.extend(f.reqs)
                if any(r.keep for r in f.reqs):
                    self.streams[k].synchronize()
                lib().nbh_batch_release(f.handle)
                self.arena.release(*f.block)
                self.busy[k] = None

    def _admit(self, k: int):
        t_start = time.perf_counter()
        reqs = []
        need = 0
        room = self.arena.largest_free
        while self.pending and len(reqs) < self.max_batch:
            nd = self._need1(self.pending[0])
            if nd > self.budget_bytes:                    # can never fit
                self.rejected.append(self.pending.popleft())
                continue
            if need + nd > room:
                break
            need += nd
            reqs.append(self.pending.popleft())
        if not reqs:
            return
        block = self.arena.alloc(need)
        assert block is not None
        off = block[0]
        rec, res_bytes, offs = self._jobs(reqs, off)
        s = self.streams[k]
        h = ctypes.c_void_p()
        rc = lib().nbh_batch_create_in(self._ptr(rec), len(reqs), ctypes.c_void_p(self.base + off + res_bytes),
                                       block[1] - res_bytes, ctypes.c_void_p(s.cuda_stream), ctypes.byref(h))
        check(rc)
        # micro-batches already overlap on the server's streams: issue each
        # one's launches in order (no per-launch fork/join host calls)
        check(lib().nbh_batch_set_streams(h, 0))
        check(lib().nbh_batch_run(h, ctypes.c_void_p(s.cuda_stream)))
        ev = torch.cuda.Event()
        ev.record(s)
        now = time.perf_counter()
        for r, o in zip(reqs, offs.tolist()):
            r.t_admit = now
            r.off = o
        self.busy[k] = _Flight(reqs, h, ev, block, k, rec)
        self.batches += 1
        self.host_admit_s += time.perf_counter() - t_start

    def result_view(self, r: Request) -> torch.Tensor:
        """The arena limbs of an in-flight request's result (valid until the
        request's micro-batch is retired)."""
        return self.mem[r.off:r.off + 8 * r.limbs].view(torch.int64)

    def poll(self) -> None:
        self._retire()
        for k in range(self.nstreams):
            if self.busy[k] is None and self.pending:
                self._admit(k)

    def drain(self) -> None:
        while not self.idle():
            self.poll()


def budget_error(req: Request) -> NbhError:
    return NbhError(NBH_ERR_BUDGET, f"request {req.rid} ({req.limbs} limbs) exceeds the HBM budget")


def poisson_arrivals(rate_per_s: float, count: int, seed: int) -> np.ndarray:
    """Arrival times (seconds from the start) of a Poisson stream."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))


def drive(srv, reqs, arrivals) -> tuple[float, float]:
    """Submit each request at its arrival time, polling in between, until
    every request has completed; returns (start, last completion) in
    perf_counter seconds.  Works with Server and NativeServer."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    i, count = 0, len(reqs)
    while i < count or not srv.idle():
        now = time.perf_counter() - t0
        while i < count and arrivals[i] <= now:
            srv.submit(reqs[i], t=t0 + arrivals[i])
            i += 1
        srv.poll()
    return t0, max(r.t_done for r in srv.done)


def latency_stats(reqs) -> dict:
    lat = np.array([r.t_done - r.t_arrival for r in reqs]) * 1e3
    if lat.size == 0:
        return {}
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()), "requests": int(lat.size)}


class NativeServer:
    """The serving loop of the library itself (nbh_server_*): queueing,
    admission by HBM budget, planning inside the arena and launches all run
    in C++; this class only maps tickets to requests.  Each request's result
    goes to its own device buffer `req.out` (int64 limbs, allocated by the
    caller before submit; at least m + n limbs)."""

    def __init__(self, budget_bytes: int, nstreams: int = 4, max_batch: int = 256, cap: int = 4096):
        init()
        self.h = ctypes.c_void_p()
        check(lib().nbh_server_create(int(budget_bytes), nstreams, max_batch, ctypes.byref(self.h)))
        self.cap = cap
        self._buf = (ctypes.c_int64 * cap)()
        self._n = ctypes.c_int64()
        self._tk = ctypes.c_int64()
        self.inflight = {}
        self.done = []
        self.rejected = []
        self.host_s = 0.0   # host time in submit and poll calls

    def submit(self, req: Request, t: float | None = None) -> Request:
        t0 = time.perf_counter()
        req.t_arrival = t0 if t is None else t
        job = self.job(req)
        rc = lib().nbh_server_submit(self.h, ctypes.byref(job), ctypes.byref(self._tk))
        if rc == NBH_ERR_BUDGET:
            self.rejected.append(req)
        else:
            check(rc)
            req.rid = self._tk.value
            self.inflight[req.rid] = req
        self.host_s += time.perf_counter() - t0
        return req

    @staticmethod
    def job(req: Request) -> nbh_job:
        """The request's nbh_job record (built once and kept on the request)."""
        job = getattr(req, "_job", None)
        if job is None:
            if req.out is None or req.out.numel() < req.limbs:
                raise ValueError("req.out must hold m + n limbs on the device")
            job = nbh_job(req.out.data_ptr(), req.dev_a, req.dev_b or None, req.m, req.n, req.bits_a, req.bits_b, 0)
            req._job = job
        return job

    def poll(self) -> None:
        t0 = time.perf_counter()
        check(lib().nbh_server_poll(self.h, self._buf, self.cap, ctypes.byref(self._n)))
        n = self._n.value
        if n:
            now = time.perf_counter()
            for i in range(n):
                r = self.inflight.pop(self._buf[i])
                r.t_done = now
                self.done.append(r)
        self.host_s += time.perf_counter() - t0

    def idle(self) -> bool:
        return not self.inflight

    def drain(self) -> None:
        check(lib().nbh_server_drain(self.h))
        while self.inflight:
            self.poll()

    @property
    def batches(self) -> int:
        b = ctypes.c_int64()
        check(lib().nbh_server_stats(self.h, None, None, ctypes.byref(b)))
        return b.value

    def close(self) -> None:
        """Destroy the queue (waits for its work); idempotent."""
        if getattr(self, "h", None):
            lib().nbh_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
