"""Autotuner for the transform schedule (which stages share an LDS tile).

    python -m nbhip.autotune [--out PATH] [--max-log2 24] [--single-max-log2 20]

For each length bucket (N <= 2^j) it times a batch of same-length products
at a representative length of the bucket under every candidate knob set
(grp_lds, mid_max, cmin, cmax), checks that each candidate's products are
bit-identical to the first candidate's, and keeps the fastest.  The result
is written as "Nmax grp_lds mid_max cmin cmax" lines; the library reads
tuning_<arch>.txt beside libnbh.so at nbh_init.  Any schedule computes the
same butterflies, so tuning changes speed only, never a result bit.

Products run alone (nbh_mul & co.) have only their own tiles to spread over
the chip, so they are tuned separately: one product per call, pipelined on
a stream, per bucket up to 2^single_max_log2; those entries are written as
"single Nmax ..." lines (longer lengths use the batch lines).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np
import torch

from . import device as D
from . import engine
from . import workloads as wl
from ._lib import PKG_DIR

CANDIDATES = [(16384, 8192, 8, 64), (8192, 4096, 8, 64), (4096, 2048, 8, 64), (4096, 2048, 8, 128),
              (4096, 1024, 8, 64), (8192, 2048, 8, 64), (16384, 4096, 8, 64), (2048, 2048, 8, 64)]


def limbs_for_length(N: int) -> int:
    """Balanced operand size (limbs) whose product uses transform length N."""
    lo, hi = 1, 1 << 26
    while lo < hi:                       # largest n with ntt length <= N
        mid = (lo + hi + 1) // 2
        if engine.fp_ntt_params(64 * mid, 64 * mid)[1] <= N:
            lo = mid
        else:
            hi = mid - 1
    return lo


def time_batch(tens, reps=7):
    """Median over `reps` runs of one planned batch (ms)."""
    B = D.Batch(tens)
    s = torch.cuda.Stream()
    B.run(s)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        B.run(s)
        e1.record(s)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    B.close()
    return float(np.median(times))


SINGLE_CANDIDATES = CANDIDATES + [(1024, 1024, 4, 32), (2048, 1024, 8, 64), (4096, 4096, 8, 64),
                                  (16384, 16384, 8, 64)]


def time_single(n, a, b, r, reps=60):
    """Median over 5 runs of `reps` pipelined single products (us per product)."""
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for _ in range(5):
            D.mul_dev(r, a, b, 64 * n, 64 * n)
        torch.cuda.synchronize()
        times = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                D.mul_dev(r, a, b, 64 * n, 64 * n)
            e1.record(s)
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) * 1e3 / reps)
    return float(np.median(times))


def tune_single(max_log2=20, log=print):
    """Per-bucket knobs for products run alone; returns [(Nmax, knobs)] with
    the buckets the batch table already serves best left out at the top."""
    engine.init()
    entries = []
    for j in range(10, max_log2 + 1):
        n = limbs_for_length(1 << j)
        N = engine.fp_ntt_params(64 * n, 64 * n)[1]
        a = D.to_dev(wl.operand("test", 900 + j, 0, n))
        b = D.to_dev(wl.operand("test", 900 + j, 1, n))
        r = torch.empty(2 * n, dtype=torch.int64, device="cuda")
        results, ref = [], None
        for cand in [None] + SINGLE_CANDIDATES:       # None: fall back to the batch table
            try:
                D.set_tuning_table([] if cand is None else [(1 << 40, cand)], single=True)
            except Exception:
                continue
            us = time_single(n, a, b, r)
            got = D.to_host(r)
            if ref is None:
                ref = got
            elif not np.array_equal(got, ref):
                raise RuntimeError(f"schedule {cand} changed the single product at N={N}")
            log(f"single N={N:>9} ({n} limbs) {cand or 'batch table'}: {us:.1f} us")
            results.append((us, cand))
        best = min(results, key=lambda x: x[0])
        entries.append((1 << j, best[1]))
        log(f"  -> single N <= 2^{j}: {best[1] or 'batch table'}")
    D.set_tuning_table([], single=True)
    # the batch table serves the top buckets by itself; inner fall-backs take the
    # batch table's knobs for that length
    while entries and entries[-1][1] is None:
        entries.pop()
    out = []
    for nmax, k in entries:
        out.append((nmax, k if k is not None else D.get_tuning_for(nmax)))
    compact = [e for i, e in enumerate(out) if not (i + 1 < len(out) and out[i + 1][1] == e[1])]
    return compact


def tune(max_log2=24, out=None, log=print, single_max_log2=0):
    engine.init()
    D.set_tuning_table([])
    info = D.device_info()
    arch = info["arch"].split(":")[0]
    entries = []
    for j in range(9, max_log2 + 1):
        n = limbs_for_length(1 << j)
        N = engine.fp_ntt_params(64 * n, 64 * n)[1]
        reps = max(1, min(2048, (1 << 25) // N))   # ~2^25 points per measurement
        a = [D.to_dev(wl.operand("test", 500 + i, 0, n)) for i in range(reps)]
        b = [D.to_dev(wl.operand("test", 500 + i, 1, n)) for i in range(reps)]
        results = []
        ref = None
        for cand in CANDIDATES:
            try:
                D.set_tuning(*cand)
            except Exception:
                continue
            tens = [(torch.empty(2 * n, dtype=torch.int64, device="cuda"), a[i], b[i]) for i in range(reps)]
            ms = time_batch(tens)
            got = [D.to_host(t[0]) for t in tens[:2]]
            if ref is None:
                ref = got
            elif not all(np.array_equal(x, y) for x, y in zip(got, ref)):
                raise RuntimeError(f"schedule {cand} changed the product at N={N}")
            log(f"N={N:>9} ({n} limbs x{reps}) {cand}: {ms:.3f} ms")
            results.append((ms, cand))
        # re-time the two fastest against each other to damp run-to-run noise
        results.sort()
        best = results[0]
        if len(results) > 1:
            again = []
            for _, cand in results[:2]:
                D.set_tuning(*cand)
                tens = [(torch.empty(2 * n, dtype=torch.int64, device="cuda"), a[i], b[i]) for i in range(reps)]
                again.append((time_batch(tens, reps=9), cand))
            best = min(again)
            log(f"  recheck: {again}")
        entries.append((1 << j, best[1]))
        log(f"  -> N <= 2^{j}: {best[1]}")
    # drop entries equal to the next one (the table is first-match by Nmax)
    compact = []
    for i, e in enumerate(entries):
        if i + 1 < len(entries) and entries[i + 1][1] == e[1]:
            continue
        compact.append(e)
    single = []
    if single_max_log2:
        D.set_tuning_table(compact)      # single products fall back to the new batch table
        single = tune_single(single_max_log2, log)
    out = out or os.path.join(PKG_DIR, f"tuning_{arch}.txt")
    with open(out, "w") as f:
        f.write(f"# autotuned on {info['arch']} ({info['cus']} CUs): Nmax grp_lds mid_max cmin cmax\n")
        for nmax, k in compact:
            f.write(f"{nmax} {k[0]} {k[1]} {k[2]} {k[3]}\n")
        for nmax, k in single:
            f.write(f"single {nmax} {k[0]} {k[1]} {k[2]} {k[3]}\n")
    log(f"wrote {out}")
    return compact, single


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--max-log2", type=int, default=24)
    ap.add_argument("--single-max-log2", type=int, default=20,
                    help="tune products run alone up to this length (0: batch table only)")
    a = ap.parse_args()
    t0 = time.time()
    tune(a.max_log2, a.out, log=lambda m: print(m, flush=True), single_max_log2=a.single_max_log2)
    print(f"autotune done in {time.time() - t0:.0f}s", file=sys.stderr)


if __name__ == "__main__":
    main()
