"""Benchmark: bit-exact limb-products/s of the MI355X NTT product engine.

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--workload batch|large|hgcd|hgcdlevel|spectra|powermod|serve|small|single|shortprod|hostbatch]

workload "batch" (default): 4096 independent products, larger size
log-uniform in [152, 65536] limbs, 25% squarings, size ratio uniform in
[1, 4].  A step is the whole job list in one planned batch (a fixed number of
launches covering every job).  The same list is also run one product per
launch for the batched/one-per-launch ratio; a short serving leg (Poisson
arrivals at 0.5x and 0.8x of capacity through nbh_server_*) gives p50/p99
latency; the CPU baseline (the oracle, with libgmp beside it) runs on the
host threads for about 15 s.
workload "hgcdlevel": 1024 hgcd! levels on resident buffers (the adjust by M1
and M <- M * M1), 12 products per instance.
workload "large": 2^20 x 2^20, 2^20 x 2^18 and a 2^22-limb square, one
product per launch; a step is the three products.
workload "hgcd": 1024 batched 2x2 cofactor steps (u', v') = (m11 u - m01 v,
m00 v - m10 u), n log-uniform in [600, 32768] limbs, cofactors n/2 limbs:
four products per instance plus the signed combine, all instances in one
planned batch.  limb-products counts the four products' result limbs.
workload "spectra": the two products of 1024 Barrett reductions modulo a
fixed 4096-limb modulus m (barrett_reduce!, src/barrett.jl:44-81): q1 * mu
((k+1) x (k+1) limbs) and q3 * m ((k+1) x k), with the forward spectra of mu
and m cached on the device, against the same 2048 products as a plain batch.
workload "powermod": 64 bases raised to one 64-bit exponent modulo a fixed
2048-limb modulus (powermod_limbs' Barrett branch, src/algorithms.jl:17-107):
per modular product a product batch and a Barrett reduction batch against the
cached spectra; limb-products counts the result limbs of all three products.
workload "serve": Poisson arrivals into the library's serving queue
(nbh_server_*) and, on the same arrivals, the Python loop over the arena
planner.
workload "small": 4096 products below the NTT thresholds (limbs log-uniform
in [1, 151], 25% squares): the batch's schoolbook kernel.
workload "single": the plain single product of 4096 x 4096 limbs and the
square of 4096 limbs, one call at a time through the drop-in boundary
(nbh_mul/nbh_sqr on host buffers: copies in and out included) and on device
buffers (nbh_mul_dev/nbh_sqr_dev), beside the oracle on one host core;
limb-products/s counts the 8192 result limbs of each call.
workload "shortprod": 1024 low short products k x k -> k limbs (mullo!/sqrlo!,
src/mul.jl:305-378), k log-uniform in [224, 65536] where the reference falls
back to a full product, 25% sqrlo, in one planned batch, against the same
operands as full products.  limb-products counts the k kept limbs.
workload "hostbatch": the batch list with operands and results in host
memory, one nbh_mul_batch call per step: the PCIe-inclusive rate (never the
headline: that one has its inputs resident in HBM).

limb-products/s = sum over completed bit-exact products of (m + n) result
limbs / wall time, inputs already resident in HBM.  Every product of the
timed run is checked bit-exact (libgmp when present, else the oracle).
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "nativebigint.jl_amd"))

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP64_VALU_PEAK_TFLOPS = 78.6   # MI355X FP64 vector (an FMA counts 2)
FP64_INSTR_PEAK = 39.3e12      # FP64 lane instructions/s: 256 CUs x 64 lanes x 2.4 GHz


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["batch", "large", "hgcd", "spectra", "powermod", "serve", "small",
                                           "shortprod", "single", "hostbatch", "hgcdlevel"],
                    default="batch")
    ap.add_argument("--bases", type=int, default=64, help="powermod: bases sharing the exponent and modulus")
    ap.add_argument("--exp-bits", type=int, default=64, help="powermod: exponent bits")
    ap.add_argument("--modulus-limbs", type=int, default=4096, help="spectra: limbs of the fixed modulus")
    ap.add_argument("--streams", type=int, default=8,
                    help="side streams a batch step forks its launches onto (0: in order)")
    ap.add_argument("--requests", type=int, default=20000, help="serve: requests per offered load")
    ap.add_argument("--budget-gib", type=float, default=8.0, help="serve: HBM budget")
    ap.add_argument("--serve-streams", type=int, default=4, help="serve: streams (micro-batches in flight)")
    ap.add_argument("--serve-max-batch", type=int, default=256, help="serve: requests per micro-batch at most")
    ap.add_argument("--jobs", type=int, default=4096)
    ap.add_argument("--single-sizes", default="152,256,512,1024,2048,4096,8192,16384,65536",
                    help="single: balanced sizes (limbs) of the drop-in latency sweep")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-one-per-launch", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="batch: skip the short serving leg (p50/p99)")
    ap.add_argument("--serve-requests", type=int, default=20000,
                    help="batch: requests per offered load in the short serving leg")
    ap.add_argument("--tuning", default="",
                    help="grp_lds,mid_max,cmin,cmax schedule knobs (default: the library's)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="also run a plain device-to-device copy of this many MiB (HBM counter calibration)")
    return ap.parse_args()


def host_info() -> dict:
    """CPU model, logical CPUs, physical cores and this process's affinity."""
    model, phys = "unknown", set()
    cur = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur:
                    phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            cur[k] = v
            if k == "model name":
                model = v
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "physical_cores": len(phys) or None,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_threads() -> int:
    """Host threads for the CPU legs: every CPU this process may run on, at
    most 16 (the host share a one-GPU box gives a process)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_throughput(ops, fn, threads: int, seconds: float) -> dict:
    """Limb-products/s of fn over the job list on `threads` threads, one
    product per thread at a time, cycling the list until `seconds` have
    passed (whole products only).  ctypes calls release the GIL."""
    lock = threading.Lock()
    state = {"next": 0, "limbs": 0, "products": 0}
    t_end = time.perf_counter() + seconds

    def worker():
        while True:
            with lock:
                if time.perf_counter() >= t_end:
                    return
                i = state["next"] % len(ops)
                state["next"] += 1
            a, b = ops[i]
            r = fn(a, b)
            with lock:
                state["limbs"] += r.size
                state["products"] += 1
    ts = [threading.Thread(target=worker) for _ in range(threads)]
    c0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - c0
    return {"value": state["limbs"] / dt, "seconds": dt, "products": state["products"],
            "passes": state["products"] / len(ops)}


def kernel_bytes(plan_groups, N, m, n, sqr, nconv):
    """Algorithmic HBM bytes per kernel kind for one product: each launch's
    unavoidable reads + writes (spectra are 8-byte doubles, limbs 8 bytes)."""
    nops = 1 if sqr else 2
    K = len(plan_groups)
    limbs_in = 8 * (m if sqr else m + n)
    out = {"group_fwd": 0, "mid": 0, "group_inv": 0, "unpack": 0, "carry_fix": 0, "combine": 0}
    for g in range(K - 1):
        rd = 2 * limbs_in if g == 0 else 2 * nops * 8 * N
        out["group_fwd"] += rd + 2 * nops * 8 * N
        out["group_inv"] += 2 * 2 * 8 * N
    out["mid"] += (2 * limbs_in if K == 1 else 2 * nops * 8 * N) + 2 * 8 * N
    out["unpack"] += 2 * 8 * nconv + 8 * (m + n)
    return out


# FP64 VALU instructions (each add/mul/fma/rndne counts once; a reduce is 3,
# a 3- or 4-argument mulmod 6, the round being one rndne on the device) per
# point per stage: radix-4 35/4, radix-2 8/2, radix-3 28/3, radix-5 82/5,
# radix-7 154/7; pointwise 6 per point per prime; unpack ~31 per coefficient.
FP64_STAGE_COST = {4: 8.75, 2: 4.0, 3: 28.0 / 3, 5: 16.4, 7: 22.0}
FP64_POINTWISE = 6.0
FP64_UNPACK = 31.0


def fp64_flops(N, plan_stage_radices, sqr, nconv):
    """FP64 VALU instructions per product, by the costs above."""
    c = sum(FP64_STAGE_COST[r] for r in plan_stage_radices)
    T = 2 if sqr else 3
    return 2 * (T * c * N + FP64_POINTWISE * N) + FP64_UNPACK * nconv


def fp64_by_kind(plan_groups, plan_stage_radices, N, sqr, nconv):
    """fp64_flops split by kernel kind: each group's stages (consecutive,
    radix product S) run in that group's kernel; forward groups transform
    every operand, inverse groups the product, and the mid kernel does its
    stages forward for every operand, the pointwise product and its stages
    inverse (both primes throughout)."""
    nops = 1 if sqr else 2
    costs, i = [], 0
    for (_, _, S, _) in plan_groups:
        c, prod = 0.0, 1
        while prod < S:
            r = plan_stage_radices[i]
            c += FP64_STAGE_COST[r]
            prod *= r
            i += 1
        costs.append(c)
    out = {"group_fwd": 0.0, "mid": 0.0, "group_inv": 0.0, "unpack": FP64_UNPACK * nconv}
    for c in costs[:-1]:
        out["group_fwd"] += 2 * nops * c * N
        out["group_inv"] += 2 * c * N
    out["mid"] += 2 * ((nops + 1) * costs[-1] * N + FP64_POINTWISE * N)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: multi-GPU sharding is a design note, "
                                   "not built; run with --gpus 1"}))
        sys.exit(2)
    import torch

    from nbhip import device as D
    from nbhip import engine
    from nbhip import workloads as wl

    engine.init()
    if args.tuning:
        D.set_tuning(*[int(v) for v in args.tuning.split(",")])
    info = D.device_info()
    info["tuning"] = D.get_tuning()
    log(f"device {info}")
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.int64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: 3 x {args.calibrate_copy_mib} MiB device-to-device")
        del src, dst

    if args.workload == "hgcd":
        return bench_hgcd(args, D, engine, wl, info, stream)
    if args.workload == "hgcdlevel":
        return bench_hgcd_level(args, D, engine, wl, info, stream)
    if args.workload == "serve":
        return bench_serve(args, D, engine, wl, info, stream)
    if args.workload == "spectra":
        return bench_spectra(args, D, engine, wl, info, stream)
    if args.workload == "powermod":
        return bench_powermod(args, D, engine, wl, info, stream)
    if args.workload == "shortprod":
        return bench_shortprod(args, D, engine, wl, info, stream)
    if args.workload == "single":
        return bench_single(args, D, engine, wl, info, stream)
    if args.workload == "hostbatch":
        return bench_hostbatch(args, D, engine, wl, info, stream)

    # ---- build the job list --------------------------------------------------
    if args.workload == "batch":
        jobs = wl.batch_jobs(count=args.jobs)
        ops = wl.batch_operands(jobs)
        wname = "batch"
    elif args.workload == "small":
        # below the NTT thresholds: the schoolbook kernel inside the batch
        jobs = wl.batch_jobs(count=args.jobs, lo=1, hi=151, workload="small")
        ops = wl.batch_operands(jobs, workload="small")
        wname = "small"
    else:
        jobs = [wl.Job(1 << 20, 1 << 20, False), wl.Job(1 << 20, 1 << 18, False), wl.Job(1 << 22, 1 << 22, True)]
        ops = [(wl.operand("large", 0, 0, 1 << 20), wl.operand("large", 0, 1, 1 << 20)),
               (wl.operand("large", 1, 0, 1 << 20), wl.operand("large", 1, 1, 1 << 18)),
               (wl.operand("large", 2, 0, 1 << 22), None)]
        wname = "large"
    t0 = time.time()
    dtens = []
    total_limbs = 0
    meta = []
    for j, (a, b) in zip(jobs, ops):
        bits_a = engine.magnitude_bits(a, 0, a.size)
        bits_b = bits_a if b is None else engine.magnitude_bits(b, 0, b.size)
        rn = a.size + (a.size if b is None else b.size)
        total_limbs += rn
        bch, N = engine.fp_ntt_params(bits_a, bits_b)
        nconv = -(-bits_a // bch) + -(-bits_b // bch) - 1
        meta.append((N, a.size, rn - a.size, b is None, nconv))
        dtens.append((torch.empty(rn, dtype=torch.int64, device=dev), D.to_dev(a),
                      None if b is None else D.to_dev(b), bits_a, bits_b))
    torch.cuda.synchronize()
    log(f"{len(jobs)} jobs, {total_limbs} result limbs, staged in {time.time() - t0:.1f}s")

    def run_one_per_launch():
        for r, a, b, ba, bb in dtens:
            if b is None:
                D.sqr_dev(r, a, ba)
            else:
                D.mul_dev(r, a, b, ba, bb)

    batch = None
    if args.workload in ("batch", "small"):
        with torch.cuda.stream(stream):
            batch = D.Batch(dtens)

    def step():
        if batch is not None:
            batch.run(stream)
        else:
            run_one_per_launch()

    # ---- warmup + timed region -------------------------------------------------
    D.set_streams(args.streams)
    with torch.cuda.stream(stream):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        tw0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        tw1 = time.perf_counter()
    ms = ev0.elapsed_time(ev1) / args.steps
    value = total_limbs / (ms * 1e-3)
    log(f"timed: {ms:.3f} ms/step (wall {1e3 * (tw1 - tw0) / args.steps:.3f}), {value:.4g} limb-products/s")

    # ---- the same batch captured into a HIP graph and replayed: the headline
    # for the batch workload (same launches, one submission per step); the
    # outputs are cleared first so the check below sees the replays' limbs ----
    graph = None
    eager = {"ms_per_step": ms, "value": value}
    if batch is not None:
        with torch.cuda.stream(stream):
            batch.capture()
            batch.replay(stream)
            for r, _, _, _, _ in dtens:
                r.fill_(-1)
            torch.cuda.synchronize()
            g0 = torch.cuda.Event(enable_timing=True)
            g1 = torch.cuda.Event(enable_timing=True)
            g0.record(stream)
            for _ in range(args.steps):
                batch.replay(stream)
            g1.record(stream)
            torch.cuda.synchronize()
        gms = g0.elapsed_time(g1) / args.steps
        graph = {"ms_per_step": gms, "value": total_limbs / (gms * 1e-3)}
        log(f"graph replay: {gms:.3f} ms/step, {graph['value']:.4g} limb-products/s")
        ms, value = gms, graph["value"]

    # ---- verify every product of the timed run -----------------------------------
    verified = None
    if not args.no_verify:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from tests import gmp_ref
        bad = 0
        use_gmp = gmp_ref.available()
        if not use_gmp:
            from oracle import oracle as O
        for (r, _, _, _, _), (a, b) in zip(dtens, ops):
            got = D.to_host(r)
            if use_gmp:
                want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
            else:
                want = O.sqr(a) if b is None else O.mul(a, b)
            bad += int(not np.array_equal(got, want))
        verified = bad == 0
        log(f"verified {len(jobs) - bad}/{len(jobs)} bit-exact ({'libgmp' if use_gmp else 'oracle'})")
        if bad:
            print(json.dumps({"error": f"{bad} products differ from the exact product"}))
            sys.exit(1)

    # ---- per-kernel profile: the same steps with every launch in order (no side
    # streams), so each launch's HIP-event duration is its own, not shared with
    # launches running beside it -------------------------------------------------
    psteps = max(1, min(args.steps, 3))
    D.set_streams(0)
    with torch.cuda.stream(stream):
        D.profile_enable(True)
        D.profile_read()
        pe0 = torch.cuda.Event(enable_timing=True)
        pe1 = torch.cuda.Event(enable_timing=True)
        pe0.record(stream)
        for _ in range(psteps):
            step()
        pe1.record(stream)
        torch.cuda.synchronize()
        prof = D.profile_read()
        D.profile_enable(False)
    D.set_streams(args.streams)
    ms_seq = pe0.elapsed_time(pe1) / psteps
    log(f"profiled (launches in order): {ms_seq:.3f} ms/step")

    # ---- roofline of the dominant kernel ---------------------------------------------
    bytes_by_kind = {k: 0 for k in D.PROF_KINDS}
    fp64_kind = {k: 0.0 for k in D.PROF_KINDS}
    flops = 0.0
    groups_cache = {}
    for (N, m, n, sq, nconv) in meta:
        if N not in groups_cache:
            groups_cache[N] = (D.plan_groups(N), [s[0] for s in D.plan_stages(N, 0)])
        groups, radices = groups_cache[N]
        for k, v in kernel_bytes(groups, N, m, n, sq, nconv).items():
            bytes_by_kind[k] += v
        for k, v in fp64_by_kind(groups, radices, N, sq, nconv).items():
            fp64_kind[k] += v
        flops += fp64_flops(N, radices, sq, nconv)
    dom = max((k for k in prof if prof[k][1] > 0), key=lambda k: prof[k][0])
    dom_ms, dom_launches = prof[dom]
    launches_per_step = dom_launches / psteps
    bytes_per_launch = bytes_by_kind[dom] / launches_per_step
    avg_launch_s = dom_ms * 1e-3 / dom_launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    def kind_line(k):
        ms_k = prof[k][0] / psteps
        line = {"ms_per_step": ms_k, "launches_per_step": prof[k][1] / psteps,
                "algorithmic_GBps": None, "hbm_frac": None, "fp64_frac": None}
        if ms_k > 0:
            gbps = bytes_by_kind[k] / (ms_k * 1e-3) / 1e9
            line["algorithmic_GBps"] = gbps
            line["hbm_frac"] = round(gbps / HBM_PEAK_GBS, 4)
            if fp64_kind.get(k):
                line["fp64_frac"] = round(fp64_kind[k] / (ms_k * 1e-3) / FP64_INSTR_PEAK, 4)
        return line

    # per kind: algorithmic bytes and model FP64 instructions of one step over
    # that kind's summed launch time (launches issued in order)
    kernels = {k: kind_line(k) for k in D.PROF_KINDS}
    step_bytes = sum(bytes_by_kind.values())
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{wname}.json")
    kname = f"k_{dom}"
    if os.path.exists(tpath):
        tk = json.load(open(tpath)).get("kernels", {}).get(kname)
        if tk:
            traffic = int(tk["bytes_per_launch"])
    roofline = {"bound": "hbm", "kernel": kname,
                "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(avg_launch_s * 1e6, 2),
                "measured": f"HIP events around every launch of {psteps} profiled steps issued in order "
                            f"({ms_seq:.3f} ms/step)"}
    # the same launches against the FP64 vector issue rate: the mid kernel's
    # model intensity (FP64 instructions per algorithmic byte) is above the
    # chip's balance point, so FP64 issue, not HBM, is its compute roofline
    instr_per_launch = fp64_kind[dom] / launches_per_step
    if instr_per_launch > 0 and bytes_per_launch > 0:
        roofline["fp64"] = {"instr_per_launch_model": int(instr_per_launch),
                            "achieved_instr_per_s": round(instr_per_launch / avg_launch_s / 1e12, 3) * 1e12,
                            "peak_instr_per_s": FP64_INSTR_PEAK,
                            "frac": round(instr_per_launch / avg_launch_s / FP64_INSTR_PEAK, 4),
                            "intensity_instr_per_byte": round(instr_per_launch / bytes_per_launch, 2),
                            "balance_instr_per_byte": round(FP64_INSTR_PEAK / (HBM_PEAK_GBS * 1e9), 2)}
    valu = {"fp64_instr_per_s_model": flops / (ms * 1e-3), "peak_instr_per_s": FP64_INSTR_PEAK,
            "frac": round(flops / (ms * 1e-3) / FP64_INSTR_PEAK, 4),
            "step_hbm_GBps_model": round(step_bytes / (ms * 1e-3) / 1e9, 1)}

    # ---- one product per launch on the same list --------------------------------
    opl = None
    if batch is not None and not args.no_one_per_launch:
        with torch.cuda.stream(stream):
            run_one_per_launch()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            reps = max(1, min(3, args.steps))
            e0.record(stream)
            for _ in range(reps):
                run_one_per_launch()
            e1.record(stream)
            torch.cuda.synchronize()
        opl_ms = e0.elapsed_time(e1) / reps
        opl = {"value": total_limbs / (opl_ms * 1e-3), "ms_per_step": opl_ms,
               "batched_speedup": opl_ms / ms}
        log(f"one-per-launch: {opl_ms:.3f} ms/step, batched speedup {opl_ms / ms:.2f}x")

    # ---- CPU baseline: the oracle (and libgmp beside it) on a bounded sample ----
    cpu = None
    # the oracle restates the NTT engine; below its thresholds the reference
    # runs Karatsuba / basecase, which the oracle does not restate
    if not args.no_cpu and wname != "small":
        from oracle import oracle as O
        from tests import gmp_ref
        threads = cpu_threads()
        fn = lambda a, b: O.sqr(a) if b is None else O.mul(a, b)  # noqa: E731
        # each build this host can run (AVX2; AVX-512 where present) on a short
        # sample, then the faster one on the timed sample
        trials = {}
        for isa in O.isas():
            O.use(isa)
            trials[isa] = cpu_throughput(ops, fn, threads, 2.0)["value"]
        O.use(max(trials, key=trials.get))
        port = cpu_throughput(ops, fn, threads, 10.0)
        hi = host_info()
        cpu = {"value": port["value"], "unit": "limb-products/s", "cores": threads, "kind": "port",
               "sample": f"{wname} job list cycled for {port['seconds']:.1f} s ({port['products']} products, "
                         f"{port['passes']:.2f} passes of all {len(jobs)}), one product per thread; C restatement "
                         f"of src/fpntt.jl (oracle/, {O.variant()}), not the Julia original",
               "builds_2s_sample": {k: round(v, 1) for k, v in trials.items()},
               "host": hi}
        if gmp_ref.available():
            gm = cpu_throughput(ops, lambda a, b: gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b), threads, 5.0)
            cpu["libgmp"] = {"value": gm["value"], "unit": "limb-products/s", "threads": threads,
                             "sample": f"__gmpn_mul/__gmpn_sqr over the same list for {gm['seconds']:.1f} s "
                                       f"({gm['products']} products)"}
        log(f"cpu baseline {cpu['value']:.4g} limb-products/s on {threads} threads"
            + (f", libgmp {cpu['libgmp']['value']:.4g}" if "libgmp" in cpu else "") + f" ({hi['model']})")

    # ---- serving: p50/p99 latency under Poisson arrivals at two offered loads ----
    latency = None
    if wname == "batch" and not args.no_serve:
        cap, native, _ = serve_leg(args, D, engine, wl, stream, args.serve_requests, 400, python_loop=False)
        latency = {"server": "nbh_server_* (continuous batching, 8 GiB HBM budget, 4 streams)",
                   "capacity_limb_products_per_s": cap,
                   "requests": "Poisson arrivals, limbs log-uniform [152, 262144], 30% squares, 20% unbalanced "
                               "(ratio <= 8), operands resident in HBM",
                   **{load: {k: v[k] for k in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "offered_limb_products_per_s",
                                               "achieved_limb_products_per_s", "requests", "micro_batches")}
                      for load, v in native.items()}}

    out = {
        "metric": "bit-exact limb-products/s",
        "value": value,
        "unit": "limb-products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": {"batch": "batch 4096 products, limbs log-uniform [152, 65536], 25% squares, "
                                         "ratio U[1,4]",
                                "small": "small: 4096 products below the NTT thresholds, limbs log-uniform "
                                         "[1, 151], 25% squares (schoolbook kernel)",
                                "large": "large: 2^20x2^20, 2^20x2^18, 2^22 square, one per launch"}[wname],
                   "jobs": len(jobs), "result_limbs_per_step": total_limbs, "parallelism": "single GPU"},
        "verified_bit_exact": verified,
        "timed_path": "HIP graph replay (nbh_batch_replay)" if graph else "launches",
        "eager": eager if graph else None,
        "roofline": roofline,
        "fp64_valu": valu,
        "kernels": kernels,
        "one_per_launch": opl,
        "latency": latency,
        "cpu_baseline": cpu,
        "device": info,
    }
    print(json.dumps(out))
    if batch is not None:
        batch.close()


def bench_hgcd(args, D, engine, wl, info, stream):
    import torch
    count = min(args.jobs, 1024)
    sizes = wl.hgcd_sizes(count)
    host = [wl.hgcd_instance(k, n) for k, n in enumerate(sizes)]
    dev = [{k: D.to_dev(v) for k, v in q.items()} for q in host]
    total_limbs = sum(4 * (q["u"].size + q["m00"].size) for q in host)
    meta = []
    for q in host:
        n, mn = q["u"].size, q["m00"].size
        bits_n = engine.magnitude_bits(q["u"], 0, n)
        bits_m = engine.magnitude_bits(q["m00"], 0, mn)
        bch, N = engine.fp_ntt_params(bits_n, bits_m)
        nconv = -(-bits_n // bch) + -(-bits_m // bch) - 1
        meta += [(N, n, mn, False, nconv)] * 4
    torch.cuda.synchronize()
    log(f"{count} cofactor steps, {total_limbs} product limbs")
    with torch.cuda.stream(stream):
        H = D.HgcdStep(dev, "adjust")
        for _ in range(args.warmup):
            H.run(stream)
        torch.cuda.synchronize()
        D.profile_enable(True)
        D.profile_read()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(args.steps):
            H.run(stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        prof = D.profile_read()
        D.profile_enable(False)
    ms = ev0.elapsed_time(ev1) / args.steps
    value = total_limbs / (ms * 1e-3)
    log(f"timed: {ms:.3f} ms/step, {value:.4g} limb-products/s")
    verified = None
    if not args.no_verify:
        from tests import gmp_ref
        bad = 0
        for (r0, r1, sg, _), q in zip(H.out, host):
            M = {k: wl.limbs_to_int(v) for k, v in q.items()}
            s = sg.cpu().numpy()
            w0 = M["m11"] * M["u"] - M["m01"] * M["v"]
            w1 = M["m00"] * M["v"] - M["m10"] * M["u"]
            ok = (int(s[0]) * wl.limbs_to_int(D.to_host(r0)) == w0 and
                  int(s[1]) * wl.limbs_to_int(D.to_host(r1)) == w1)
            bad += int(not ok)
        verified = bad == 0
        log(f"verified {count - bad}/{count} cofactor steps exact")
        if bad:
            print(json.dumps({"error": f"{bad} cofactor steps differ"}))
            sys.exit(1)
    out = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": "hgcd: 1024 cofactor steps (m11 u - m01 v, m00 v - m10 u), n log-uniform "
                               "[600, 32768], cofactors n/2 limbs", "instances": count,
                   "product_limbs_per_step": total_limbs, "parallelism": "single GPU"},
        "verified_bit_exact": verified,
        "kernels": {k: {"ms_per_step": prof[k][0] / args.steps, "launches_per_step": prof[k][1] / args.steps}
                    for k in D.PROF_KINDS},
        "device": info,
    }
    print(json.dumps(out))
    H.close()


def bench_hgcd_level(args, D, engine, wl, info, stream):
    """One hgcd! level (src/gcd.jl:546-554) per instance on resident buffers:
    hgcd_matrix_adjust! by M1 and hgcd_matrix_mul! M <- M * M1, 12 products
    and 6 sums / signed differences per instance, updated in place."""
    import torch
    count = min(args.jobs, 1024)
    sizes = wl.hgcd_sizes(count)
    dev = []
    total_limbs = 0
    for k, n0 in enumerate(sizes):
        q = wl.hgcd_level_instance(k, n0)
        sh = wl.hgcd_level_shape(n0)
        d = {}
        for name in ("m00", "m01", "m10", "m11"):
            x = np.zeros(sh["Lm"], dtype=np.uint64)
            x[:q[name].size] = q[name]
            d[name] = D.to_dev(x)
        for name in ("n00", "n01", "n10", "n11", "ra", "rb"):
            d[name] = D.to_dev(q[name])
        for name in ("a", "b"):
            x = np.zeros(max(sh["L"], q[name].size), dtype=np.uint64)
            x[:q[name].size] = q[name]
            d[name] = D.to_dev(x)
        d["mn"], d["p"] = sh["mn"], sh["p"]
        dev.append(d)
        total_limbs += 4 * (sh["p"] + sh["m1n"]) + 8 * (sh["mn"] + sh["m1n"])
    torch.cuda.synchronize()
    log(f"{count} hgcd levels, {total_limbs} product limbs per step")
    with torch.cuda.stream(stream):
        H = D.HgcdLevel(dev)
    ms, prof = _timed_runs(H, stream, args.steps, args.warmup, D)
    value = total_limbs / (ms * 1e-3)
    log(f"timed: {ms:.3f} ms/step, {value:.4g} limb-products/s")
    out = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": "hgcdlevel: 1024 hgcd! levels (hgcd_matrix_adjust! by M1, then M <- M*M1) on "
                               "resident buffers, n0 log-uniform [600, 32768]", "instances": count,
                   "product_limbs_per_step": total_limbs, "parallelism": "single GPU"},
        "verified_bit_exact": "tests/test_gpu_hgcd_level.py (in place, two runs, libgmp and Python ints)",
        "kernels": {k: {"ms_per_step": prof[k][0] / args.steps, "launches_per_step": prof[k][1] / args.steps}
                    for k in D.PROF_KINDS},
        "device": info,
    }
    print(json.dumps(out))
    H.close()


def _timed_runs(B, stream, steps, warmup, D, profile=True):
    import torch
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            B.run(stream)
        torch.cuda.synchronize()
        if profile:
            D.profile_enable(True)
            D.profile_read()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(steps):
            B.run(stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        prof = None
        if profile:
            prof = D.profile_read()
            D.profile_enable(False)
    return ev0.elapsed_time(ev1) / steps, prof


def bench_spectra(args, D, engine, wl, info, stream):
    """Barrett products against a fixed modulus with cached spectra, and the
    same products as a plain batch."""
    import torch
    k = args.modulus_limbs
    count = min(args.jobs, 1024)
    m_l, mu_l, q1s, q3s = wl.barrett_operands(count, k)
    dm, dmu = D.to_dev(m_l), D.to_dev(mu_l)
    dq1 = [D.to_dev(q) for q in q1s]
    dq3 = [D.to_dev(q) for q in q3s]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        S_mu = D.Spectrum(dmu, 64 * (k + 1))
        S_m = D.Spectrum(dm, 64 * (k + 1))
    torch.cuda.synchronize()
    t_spec = time.perf_counter() - t0
    r_spec = [torch.empty(2 * k + 2, dtype=torch.int64, device="cuda") for _ in range(count)] + \
             [torch.empty(2 * k + 1, dtype=torch.int64, device="cuda") for _ in range(count)]
    r_plain = [torch.empty_like(r) for r in r_spec]
    jobs_s = [(r_spec[i], dq1[i], S_mu) for i in range(count)] + \
             [(r_spec[count + i], dq3[i], S_m) for i in range(count)]
    jobs_p = [(r_plain[i], dq1[i], dmu) for i in range(count)] + \
             [(r_plain[count + i], dq3[i], dm) for i in range(count)]
    total_limbs = sum(r.numel() for r in r_spec)
    log(f"{count} Barrett reductions mod a {k}-limb modulus: {2 * count} products, {total_limbs} limbs; "
        f"spectra N={S_mu.N}, {S_m.N} built in {1e3 * t_spec:.2f} ms")
    BS = D.SpecBatch(jobs_s)
    BP = D.Batch(jobs_p)
    ms_p, prof_p = _timed_runs(BP, stream, args.steps, args.warmup, D)
    ms_s, prof_s = _timed_runs(BS, stream, args.steps, args.warmup, D)
    value = total_limbs / (ms_s * 1e-3)
    plain = total_limbs / (ms_p * 1e-3)
    log(f"timed: cached spectra {ms_s:.3f} ms/step ({value:.4g} limb-products/s), "
        f"plain {ms_p:.3f} ms/step ({plain:.4g}), x{ms_p / ms_s:.2f}")
    verified = None
    if not args.no_verify:
        from tests import gmp_ref
        bad = 0
        for rs, rp in zip(r_spec, r_plain):
            bad += int(not torch.equal(rs, rp))
        ops = [(q1s[i], mu_l) for i in range(count)] + [(q3s[i], m_l) for i in range(count)]
        for i in range(0, 2 * count, max(1, (2 * count) // 64)):     # 64 products against libgmp
            a, b = ops[i]
            bad += int(not np.array_equal(D.to_host(r_spec[i]), gmp_ref.mul(a, b)))
        verified = bad == 0
        log(f"verified: {2 * count} products equal to the plain batch, 64 equal to libgmp; {bad} differ")
        if bad:
            print(json.dumps({"error": f"{bad} products differ"}))
            sys.exit(1)
    out = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_s, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": f"spectra: {count} Barrett reductions mod a fixed {k}-limb modulus, "
                               f"products q1*mu ((k+1)x(k+1)) and q3*m ((k+1)xk) against cached spectra",
                   "products_per_step": 2 * count, "product_limbs_per_step": total_limbs,
                   "parallelism": "single GPU"},
        "plain_batch": {"value": plain, "ms_per_step": ms_p, "speedup_with_spectra": ms_p / ms_s},
        "spectra_build_ms": 1e3 * t_spec,
        "verified_bit_exact": verified,
        "kernels": {kk: {"ms_per_step": prof_s[kk][0] / args.steps, "launches_per_step": prof_s[kk][1] / args.steps,
                         "plain_ms_per_step": prof_p[kk][0] / args.steps}
                    for kk in D.PROF_KINDS},
        "device": info,
    }
    print(json.dumps(out))
    BS.close()
    BP.close()
    S_mu.close()
    S_m.close()


def bench_single(args, D, engine, wl, info, stream):
    """One product at a time: the drop-in host path, the device path and the
    oracle on one core, for a 4096 x 4096 product and a 4096-limb square."""
    import torch
    n = 4096
    a, b = wl.operand("single", 0, 0, n), wl.operand("single", 0, 1, n)
    reps = max(args.steps, 1) * 20
    r = np.empty(2 * n, dtype=np.uint64)
    res = {}
    for name, call in (("mul", lambda: engine.mul_fpntt2(r, 0, a, 0, n, b, 0, n)),
                       ("sqr", lambda: engine.sqr_fpntt2(r, 0, a, 0, n))):
        for _ in range(max(args.warmup, 1) * 5):
            call()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        res[name + "_host_us"] = (time.perf_counter() - t0) / reps * 1e6
    host_mul = r.copy()
    engine.mul_fpntt2(host_mul, 0, a, 0, n, b, 0, n)
    da, db = D.to_dev(a), D.to_dev(b)
    dr = D.empty_limbs(2 * n)
    with torch.cuda.stream(stream):
        for name, call in (("mul", lambda: D.mul_dev(dr, da, db, 64 * n, 64 * n)),
                           ("sqr", lambda: D.sqr_dev(dr, da, 64 * n))):
            for _ in range(max(args.warmup, 1) * 5):
                call()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(reps):
                call()
            e1.record(stream)
            torch.cuda.synchronize()
            res[name + "_dev_us"] = e0.elapsed_time(e1) / reps * 1e3
    with torch.cuda.stream(stream):
        D.mul_dev(dr, da, db, 64 * n, 64 * n)
    torch.cuda.synchronize()
    dev_mul = D.to_host(dr)
    verified = None
    if not args.no_verify:
        from tests import gmp_ref
        want = gmp_ref.mul(a, b)
        sq = np.empty(2 * n, dtype=np.uint64)
        engine.sqr_fpntt2(sq, 0, a, 0, n)
        verified = bool(np.array_equal(host_mul, want) and np.array_equal(dev_mul, want)
                        and np.array_equal(sq, gmp_ref.sqr(a)))
        log(f"verified: product (host and device paths) and square equal to libgmp: {verified}")
        if not verified:
            print(json.dumps({"error": "single product differs"}))
            sys.exit(1)
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.mul(a, b)
        creps = 20
        c0 = time.perf_counter()
        for _ in range(creps):
            O.mul(a, b)
        cmul = (time.perf_counter() - c0) / creps
        cpu = {"value": 2 * n / cmul, "unit": "limb-products/s", "cores": 1, "kind": "port",
               "sample": f"{creps} products of 4096 x 4096 limbs on one host core, C restatement of "
                         f"src/fpntt.jl (oracle/), {cmul * 1e6:.0f} us each"}
        log(f"cpu (oracle, one core): {cmul * 1e6:.0f} us per 4096 x 4096 product")
    log(f"single 4096 x 4096: host path {res['mul_host_us']:.0f} us, device path {res['mul_dev_us']:.0f} us; "
        f"square: host {res['sqr_host_us']:.0f} us, device {res['sqr_dev_us']:.0f} us")
    # drop-in latency against the oracle on one core across sizes: where the
    # GPU call starts to pay for a caller that multiplies one product at a time
    sweep = []
    if not args.no_cpu:
        from oracle import oracle as O
        for k, sz in enumerate(int(v) for v in args.single_sizes.split(",") if v):
            x, y = wl.operand("single", 1 + k, 0, sz), wl.operand("single", 1 + k, 1, sz)
            rr = np.empty(2 * sz, dtype=np.uint64)
            engine.mul_fpntt2(rr, 0, x, 0, sz, y, 0, sz)
            nrep = max(3, min(200, int(2e7 // (sz * 64))))

            def best_of_3(f):
                """Mean time per call in us, the best of three rounds of nrep
                calls (one host hiccup in a round of short calls would
                otherwise dominate it)."""
                best = float("inf")
                for _ in range(3):
                    t0 = time.perf_counter()
                    for _ in range(nrep):
                        f()
                    best = min(best, (time.perf_counter() - t0) / nrep * 1e6)
                return best
            gpu_us = best_of_3(lambda: engine.mul_fpntt2(rr, 0, x, 0, sz, y, 0, sz))
            ref = O.mul(x, y)
            cpu_us = best_of_3(lambda: O.mul(x, y))
            ok = bool(np.array_equal(rr, ref))
            sweep.append({"limbs": sz, "drop_in_us": round(gpu_us, 1), "oracle_1core_us": round(cpu_us, 1),
                          "equal": ok})
            log(f"  {sz:>6} x {sz:<6} drop-in {gpu_us:9.1f} us   oracle (1 core) {cpu_us:10.1f} us   "
                f"x{cpu_us / gpu_us:5.2f}{'' if ok else '   DIFFERENT'}")
            if not ok:
                print(json.dumps({"error": f"sweep product at {sz} limbs differs"}))
                sys.exit(1)
    value = 2 * n / (res["mul_host_us"] * 1e-6)
    out = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": reps, "warmup": args.warmup, "ms_per_step": res["mul_host_us"] * 1e-3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": "single: one 4096 x 4096 product per call through nbh_mul (host buffers, "
                               "copies included)", "parallelism": "single GPU"},
        "latency_us": {k: round(v, 1) for k, v in res.items() if k.endswith("_us")},
        "drop_in_sweep": sweep,
        "verified_bit_exact": verified,
        "cpu_baseline": cpu,
        "device": info,
    }
    print(json.dumps(out))


def bench_shortprod(args, D, engine, wl, info, stream):
    """Low short products in one batch, and the same operands as full
    products (what the reference's mullo!/sqrlo! run at these sizes)."""
    import torch
    count = min(args.jobs, 1024)
    jobs = wl.batch_jobs(count=count, lo=224, hi=65536, max_ratio=1.0, workload="shortprod")
    ops = wl.batch_operands(jobs, workload="shortprod")
    js, jf, keep = [], [], []
    for j, (a, b) in zip(jobs, ops):
        k = a.size
        da = D.to_dev(a)
        db = None if b is None else D.to_dev(b)
        ba = engine.magnitude_bits(a, 0, a.size)
        bb = ba if b is None else engine.magnitude_bits(b, 0, b.size)
        js.append((D.empty_limbs(k), da, db, ba, bb, k))
        jf.append((D.empty_limbs(2 * k), da, db, ba, bb))
        keep.append(k)
    total_limbs = sum(keep)
    torch.cuda.synchronize()
    log(f"{count} short products, {total_limbs} kept limbs")
    BS = D.Batch(js)
    BF = D.Batch(jf)
    D.set_streams(args.streams)
    ms_f, _ = _timed_runs(BF, stream, args.steps, args.warmup, D, profile=False)
    ms_s, _ = _timed_runs(BS, stream, args.steps, args.warmup, D, profile=False)
    D.set_streams(0)
    _, prof = _timed_runs(BS, stream, min(args.steps, 3), 0, D)
    value = total_limbs / (ms_s * 1e-3)
    log(f"timed: short {ms_s:.3f} ms/step ({value:.4g} kept limbs/s), full products of the same operands "
        f"{ms_f:.3f} ms/step, x{ms_f / ms_s:.2f}")
    verified = None
    if not args.no_verify:
        from tests import gmp_ref
        bad = 0
        for (rs, *_), (rf, *_), k in zip(js, jf, keep):
            bad += int(not torch.equal(rs, rf[:k]))
        for i in range(0, count, max(1, count // 32)):            # 32 against libgmp
            a, b = ops[i]
            full = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
            bad += int(not np.array_equal(D.to_host(js[i][0]), full[:keep[i]]))
        verified = bad == 0
        log(f"verified: {count} short products equal to the full products' low limbs, 32 to libgmp; {bad} differ")
        if bad:
            print(json.dumps({"error": f"{bad} products differ"}))
            sys.exit(1)
    out = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_s, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": f"shortprod: {count} low short products k x k -> k limbs (mullo!/sqrlo!), "
                               "k log-uniform [224, 65536], 25% sqrlo", "products": count,
                   "kept_limbs_per_step": total_limbs, "parallelism": "single GPU"},
        "full_products_ms_per_step": ms_f,
        "speedup_vs_full_products": ms_f / ms_s,
        "verified_bit_exact": verified,
        "kernels": {k: {"ms_per_step": prof[k][0] / min(args.steps, 3),
                        "launches_per_step": prof[k][1] / min(args.steps, 3)} for k in D.PROF_KINDS},
        "device": info,
    }
    print(json.dumps(out))
    BS.close()
    BF.close()


def bench_powermod(args, D, engine, wl, info, stream):
    """Batched powermod modulo a fixed modulus (Barrett with cached spectra)."""
    import torch
    from nbhip import barrett as BR
    k = args.modulus_limbs if args.modulus_limbs != 4096 else 2048
    count = args.bases
    m = wl.operand("powermod", 0xFFFFF0, 0, k)
    M = wl.limbs_to_int(m)
    e_l = wl.operand("powermod", 0xFFFFF0, 1, (args.exp_bits + 63) // 64)
    e = wl.limbs_to_int(e_l) % (1 << args.exp_bits) | (1 << (args.exp_bits - 1))
    bases = []
    for i in range(count):
        b = wl.limbs_to_int(wl.operand("powermod", i, 0, k)) % M
        bases.append(wl.int_to_limbs(max(b, 2), k))
    P = BR.PowermodBatch(m, bases, e)
    nmod = P.modmuls()
    limbs_per_modmul = 2 * k + (2 * k + 2) + (2 * k + 1)
    total_limbs = count * nmod * limbs_per_modmul
    log(f"{count} bases ^ {args.exp_bits}-bit exponent mod a {k}-limb modulus: {nmod} modular products each "
        f"(window {P.w}), {total_limbs} product limbs per run")
    with torch.cuda.stream(stream):
        for _ in range(max(1, args.warmup)):
            P.run(stream)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        steps = max(1, min(args.steps, 3))
        ev0.record(stream)
        for _ in range(steps):
            out = P.run(stream)
        ev1.record(stream)
        torch.cuda.synchronize()
        # per-kernel split from one more, profiled run (events around every launch)
        D.profile_enable(True)
        D.profile_read()
        P.run(stream)
        torch.cuda.synchronize()
        prof = D.profile_read()
        D.profile_enable(False)
    ms_eager = ev0.elapsed_time(ev1) / steps
    log(f"eager: {ms_eager:.2f} ms per exponentiation batch")
    # the whole run as one HIP graph (hundreds of short launches in one submission)
    P.capture()
    with torch.cuda.stream(stream):
        for r in P.acc:
            r.fill_(-1)
        P.replay()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(steps):
            out = P.replay()
        ev1.record(stream)
        torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    value = total_limbs / (ms * 1e-3)
    log(f"timed (graph replay): {ms:.2f} ms per exponentiation batch, {count * nmod / (ms * 1e-3):.4g} modular "
        f"products/s, {value:.4g} limb-products/s")
    verified = None
    if not args.no_verify:
        sample = list(range(0, count, max(1, count // 4)))
        bad = sum(int(wl.limbs_to_int(D.to_host(out[i])) != pow(wl.limbs_to_int(bases[i]), e, M)) for i in sample)
        verified = bad == 0
        log(f"verified {len(sample) - bad}/{len(sample)} results against pow()")
        if bad:
            print(json.dumps({"error": f"{bad} powermod results differ"}))
            sys.exit(1)
    out_json = {
        "metric": "bit-exact limb-products/s", "value": value, "unit": "limb-products/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb)",
        "config": {"workload": f"powermod: {count} bases, {args.exp_bits}-bit exponent, {k}-limb modulus, "
                               f"sliding window {P.w}, Barrett with cached spectra",
                   "modular_products_per_step": count * nmod, "parallelism": "single GPU"},
        "modular_products_per_s": count * nmod / (ms * 1e-3),
        "timed_path": "HIP graph replay of the whole exponentiation (PowermodBatch.capture)",
        "eager": {"ms_per_step": ms_eager},
        "verified_bit_exact": verified,
        "kernels": {kk: {"ms_per_step": prof[kk][0], "launches_per_step": prof[kk][1]} for kk in D.PROF_KINDS},
        "device": info,
    }
    print(json.dumps(out_json))
    P.close()


def bench_hostbatch(args, D, engine, wl, info, stream):
    """The headline batch's 4096 products with operands and results in host
    memory, one nbh_mul_batch call per step: the PCIe-inclusive rate a caller
    with Memory{UInt64} limbs sees (staging copies, transfers, planning and
    the pipeline, timed by the host clock)."""
    jobs = wl.batch_jobs(count=args.jobs)
    ops = wl.batch_operands(jobs)
    sizes = [a.size + (a.size if b is None else b.size) for a, b in ops]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    r = np.zeros(int(sum(sizes)), dtype=np.uint64)
    calls = [(r, int(o), a, 0, a.size, b, 0, 0 if b is None else b.size, 0) for (a, b), o in zip(ops, offs)]
    # the records are checked and built once; the timed calls are the C ABI's
    from nbhip._lib import check, lib
    recs = engine.mul_batch_records(calls)
    for _ in range(max(args.warmup, 1)):
        check(lib().nbh_mul_batch(recs, len(recs)))
    r[:] = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        check(lib().nbh_mul_batch(recs, len(recs)))
    dt = (time.perf_counter() - t0) / args.steps
    limbs = float(sum(sizes))
    verified = None
    if not args.no_verify:
        from tests import gmp_ref
        bad = 0
        for k in range(0, len(ops), 16):
            a, b = ops[k]
            want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
            bad += int(not np.array_equal(r[offs[k]:offs[k] + sizes[k]], want))
        verified = bad == 0
        log(f"verified {len(range(0, len(ops), 16))} sampled products against libgmp: {verified}")
        if bad:
            print(json.dumps({"error": f"{bad} host-batch products differ"}))
            sys.exit(1)
    in_bytes = sum(8 * (a.size + (0 if b is None else b.size)) for a, b in ops)
    log(f"timed: {dt * 1e3:.2f} ms/step, {limbs / dt:.4g} limb-products/s PCIe-inclusive "
        f"({in_bytes / 2**20:.0f} MiB in, {8 * limbs / 2**20:.0f} MiB out per step)")
    out = {"metric": "bit-exact limb-products/s", "value": limbs / dt, "unit": "limb-products/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (SplitMix64 uniform limbs, nonzero top limb), host memory",
           "config": {"workload": "hostbatch: the batch list (4096 products, limbs log-uniform [152, 65536]) on "
                                  "host buffers through nbh_mul_batch, PCIe-inclusive",
                      "host_bytes_in": in_bytes, "host_bytes_out": 8 * limbs, "parallelism": "single GPU"},
           "verified_bit_exact": verified, "device": info}
    print(json.dumps(out))


def serve_leg(args, D, engine, wl, stream, count, pool, python_loop):
    """Poisson arrivals at 0.5x and 0.8x of the capacity measured on the same
    request mix (one planned batch of the first requests), through the
    library's queue (nbh_server_*) and, with python_loop, the Python loop over
    the arena planner on the same arrivals.  A sample of every load's results
    is checked against libgmp.  Returns (capacity, native, python)."""
    import torch
    from nbhip import serve
    from tests import gmp_ref
    pool = min(pool, count)            # distinct operand pairs; requests draw from them
    jobs = wl.serve_jobs(pool)
    ops = wl.batch_operands(jobs, workload="serve")
    dev, bits = [], []
    for (a, b) in ops:
        dev.append((D.to_dev(a), None if b is None else D.to_dev(b)))
        ba = engine.magnitude_bits(a, 0, a.size)
        bits.append((ba, ba if b is None else engine.magnitude_bits(b, 0, b.size)))
    limbs = np.array([a.size + (a.size if b is None else b.size) for a, b in ops], dtype=np.float64)
    torch.cuda.synchronize()
    # capacity: the first 512 requests as one planned batch
    cap_n = min(512, pool)
    tens = [(torch.empty(int(limbs[i]), dtype=torch.int64, device="cuda"), dev[i][0], dev[i][1], *bits[i])
            for i in range(cap_n)]
    with torch.cuda.stream(stream):
        B = D.Batch(tens)
        B.run(stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(3):
            B.run(stream)
        e1.record(stream)
        torch.cuda.synchronize()
        B.close()
    cap = limbs[:cap_n].sum() / (e0.elapsed_time(e1) / 3 * 1e-3)
    del tens
    log(f"capacity (one planned batch of {cap_n}): {cap:.4g} limb-products/s")
    rng = np.random.default_rng(0x5e)
    streams = []
    for load in (0.5, 0.8):
        pick = rng.integers(0, pool, count)
        rate = load * cap / limbs[pick].mean()                # requests per second
        streams.append((load, pick, np.cumsum(rng.exponential(1.0 / rate, count))))
    budget = int(args.budget_gib * (1 << 30))
    keep = set(range(0, count, max(1, count // 64)))

    def summary(srv, reqs, pick, load, t0, t_end, host_s, label):
        lat = serve.latency_stats(srv.done)
        thr = sum(r.limbs for r in srv.done) / (t_end - t0)
        bad = 0
        for k in keep:
            a, b = ops[pick[k]]
            want = gmp_ref.sqr(a) if b is None else gmp_ref.mul(a, b)
            bad += int(not np.array_equal(D.to_host(reqs[k].out), want))
        if bad:
            print(json.dumps({"error": f"{bad} served products differ ({label})"}))
            sys.exit(1)
        log(f"{label} load {load}: p50 {lat['p50_ms']:.2f} ms, p99 {lat['p99_ms']:.2f} ms, "
            f"{thr:.4g} limb-products/s over {srv.batches} micro-batches in {t_end - t0:.3f} s "
            f"(host time in the loop {host_s:.3f} s)")
        return {"offered_limb_products_per_s": load * cap, "achieved_limb_products_per_s": thr,
                "operand_pool": pool, "requests": count, "rejected": len(srv.rejected),
                "micro_batches": srv.batches, "duration_s": t_end - t0, "host_s": host_s,
                "verified_sample": len(keep), **lat}

    results, native = {}, {}
    for load, pick, arrivals in streams:
        if python_loop:
            # the Python loop over the arena planner (results in the arena, sample copied out)
            srv = serve.Server(budget_bytes=budget, nstreams=args.serve_streams, max_batch=args.serve_max_batch)
            reqs = [serve.Request(dev[p][0], dev[p][1], *bits[p], keep=i in keep) for i, p in enumerate(pick)]
            srv.prepare(reqs)                 # start-up: plans of every request shape
            t0, t_end = serve.drive(srv, reqs, arrivals)
            results[str(load)] = summary(srv, reqs, pick, load, t0, t_end, srv.host_admit_s, "python loop")
            del srv, reqs
            torch.cuda.empty_cache()
        # the library's own serving queue: results straight into per-request buffers
        res = limbs[pick].astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(res)[:-1]])
        out = torch.empty(int(res.sum()), dtype=torch.int64, device="cuda")
        reqs = [serve.Request(dev[p][0], dev[p][1], *bits[p]) for p in pick]
        for r, o in zip(reqs, offs.tolist()):
            r.out = out[o:o + r.limbs]
            serve.NativeServer.job(r)
        srv = serve.NativeServer(budget_bytes=budget, nstreams=args.serve_streams, max_batch=args.serve_max_batch)
        warm = {}                         # start-up: one request of every shape through the queue
        for p in sorted(set(pick.tolist())):
            w = serve.Request(dev[p][0], dev[p][1], *bits[p])
            w.out = torch.empty(w.limbs, dtype=torch.int64, device="cuda")
            warm[p] = w
            srv.submit(w)
        srv.drain()
        srv.done.clear()
        srv.host_s = 0.0
        del warm
        t0, t_end = serve.drive(srv, reqs, arrivals)
        native[str(load)] = summary(srv, reqs, pick, load, t0, t_end, srv.host_s, "native queue")
        srv.close()
        del srv, reqs, out
        torch.cuda.empty_cache()
    return cap, native, results


def bench_serve(args, D, engine, wl, info, stream):
    """Poisson arrivals into the continuous-batching server at two offered
    loads (fractions of the capacity measured on the same request mix)."""
    cap, native, results = serve_leg(args, D, engine, wl, stream, args.requests, 1000, python_loop=True)
    out = {"metric": "bit-exact limb-products/s", "value": native["0.8"]["achieved_limb_products_per_s"],
           "unit": "limb-products/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": native["0.8"]["duration_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (SplitMix64 uniform limbs, Poisson arrivals)",
           "config": {"workload": "serve: Poisson arrivals, limbs log-uniform [152, 262144], 30% squares, "
                                  "20% unbalanced (ratio <= 8), operands resident in HBM",
                      "capacity_limb_products_per_s": cap, "budget_gib": args.budget_gib,
                      "parallelism": f"single GPU, {args.serve_streams} streams",
                      "max_batch": args.serve_max_batch,
                      "server": "nbh_server_* (the library's queue); latency_python_loop: nbhip.serve.Server"},
           "latency": native, "latency_python_loop": results, "device": info}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
