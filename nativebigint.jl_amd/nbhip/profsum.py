"""Summarise a rocprofv3 results database into a short text table.

    python -m nbhip.profsum <results.db> [--title T] [--command C] > summary.md

Kernel-trace runs (--kernel-trace --stats) give calls, total and average
duration per kernel; counter runs (--pmc ...) give each counter summed per
kernel, and per dispatch.  Kernel names are shortened to the function name.
"""
from __future__ import annotations

import argparse
import collections
import re
import sqlite3


def short(name: str) -> str:
    m = re.match(r"(?:void\s+)?([\w:]+(?:<[^>]*>)?)\(", name)
    return m.group(1) if m else name[:60]


def tables(cur):
    return {r[0] for r in cur.execute("select name from sqlite_master where type in ('table','view')")}


def kernel_stats(cur):
    rows = cur.execute("select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
                       "from kernels group by name order by sum(duration) desc").fetchall()
    total = sum(r[2] for r in rows) or 1
    out = ["| kernel | calls | total us | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for name, n, tot, avg, mn, mx in rows:
        out.append(f"| {short(name)} | {n} | {tot / 1e3:.1f} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} "
                   f"| {100 * tot / total:.1f} |")
    return out


def counter_stats(cur):
    rows = cur.execute("select kernel_name, dispatch_id, counter_name, sum(value) from counters_collection "
                       "group by kernel_name, dispatch_id, counter_name").fetchall()
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    disp = collections.defaultdict(set)
    for name, d, c, v in rows:
        agg[short(name)][c] += v
        disp[short(name)].add(d)
    counters = sorted({c for k in agg for c in agg[k]})
    out = ["| kernel | dispatches | " + " | ".join(counters) + " |",
           "|---|---|" + "---|" * len(counters)]
    for k in sorted(agg, key=lambda k: -max(agg[k].values())):
        out.append(f"| {k} | {len(disp[k])} | " + " | ".join(f"{agg[k][c]:.4g}" for c in counters) + " |")
    out.append("")
    out.append("per dispatch (sum / dispatches):")
    out.append("")
    out.append("| kernel | " + " | ".join(counters) + " |")
    out.append("|---|" + "---|" * len(counters))
    for k in sorted(agg, key=lambda k: -max(agg[k].values())):
        n = max(1, len(disp[k]))
        out.append(f"| {k} | " + " | ".join(f"{agg[k][c] / n:.4g}" for c in counters) + " |")
    return out


def counter_per_dispatch(db: str, counter: str):
    """{kernel short name: [per-dispatch sums of `counter`]}."""
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                       "where counter_name = ? group by kernel_name, dispatch_id", (counter,)).fetchall()
    out = collections.defaultdict(list)
    for name, _, v in rows:
        out[short(name)].append(v)
    return out


def _names(full: str, k: str) -> bool:
    """Whether a demangled kernel name is kernel k itself (k_group_fwd does
    not name k_group_fwd_w6)."""
    return re.search(r"(^|::|\s)" + re.escape(k) + r"(?![\w])", full) is not None


def traffic(fetch_db: str, write_db: str, kernels, copy_bytes: int, copy_kernel: str = "copyBuffer"):
    """HBM bytes per launch of each kernel from FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch), each scaled by the factor that makes the calibration
    copy of `copy_bytes` read and write exactly that many bytes."""
    f = counter_per_dispatch(fetch_db, "FETCH_SIZE")
    w = counter_per_dispatch(write_db, "WRITE_SIZE")
    cf = [v for k, vs in f.items() if copy_kernel in k for v in vs]
    cw = [v for k, vs in w.items() if copy_kernel in k for v in vs]
    big_f, big_w = max(cf), max(cw)                 # the calibration copy is the largest copy
    f_scale = copy_bytes / (big_f * 1024)
    w_scale = copy_bytes / (big_w * 1024)
    out = {"copy_bytes": copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for k in kernels:
        fk = [v for name, vs in f.items() if _names(name, k) for v in vs]
        wk = [v for name, vs in w.items() if _names(name, k) for v in vs]
        if not fk or not wk:
            continue
        fetch = sum(fk) * 1024 * f_scale / len(fk)
        write = sum(wk) * 1024 * w_scale / len(wk)
        out["kernels"][k] = {"dispatches": len(fk), "fetch_bytes_per_launch": fetch,
                             "write_bytes_per_launch": write, "bytes_per_launch": fetch + write}
    return out


def _md_table(path: str) -> dict:
    """The per-dispatch table of a counter summary written by this tool."""
    txt = open(path).read().split("per dispatch")[1]
    lines = txt.split("\n")
    hdr = [x for x in lines if x.startswith("| kernel")][0]
    cols = [c.strip() for c in hdr.strip("|").split("|")]
    out = {}
    for row in (x for x in lines if x.startswith("| nbh::")):
        v = [c.strip() for c in row.strip("|").split("|")]
        out[v[0]] = {cols[i]: float(v[i]) for i in range(1, len(v))}
    return out


def counters_report(pass1: str, pass2: str, trace: str, commands: list[str], kernels: list[str],
                    cus: int = 256, ghz: float = 2.4) -> str:
    """Derived SQ-counter figures per kernel from two counter summaries of
    this tool (pass1: SQ_INSTS_VALU, its F64 parts, SQ_INSTS_SALU,
    SQ_INSTS_LDS; pass2: SQ_WAVE_CYCLES) and the kernel-trace summary of the
    same command (average launch time)."""
    t1, t2 = _md_table(pass1), _md_table(pass2)
    avg = {}
    for line in open(trace):
        m = re.match(r"\| (nbh::\S+) \| (\d+) \| [\d.]+ \| ([\d.]+) \|", line)
        if m:
            avg[m.group(1)] = float(m.group(3))
    out = ["# SQ counters: batch workload, launches in order", "",
           "Commands (one pass each, summarised per dispatch by `python3 -m nbhip.profsum <results.db>`, "
           "then by `python3 -m nbhip.profsum --counters pass1.md pass2.md kernel_trace_batch.md`):", "",
           "```", *commands, "```", "",
           "Average launch time from `kernel_trace_batch.md` (same workload and flags). Derived columns:",
           "VALU wave-instructions per dispatch; FP64 share of them; VALU issue use = VALU wave-instructions",
           f"/ ({cus} CUs x 4 SIMDs x launch cycles at {ghz} GHz / 4 cycles per wave64 instruction); resident "
           "waves", f"per CU = SQ_WAVE_CYCLES (quad-cycles) x 4 / launch cycles / {cus}.", "",
           "| kernel | avg us | VALU instr | FP64 share | VALU issue use | FP64 issue use | SALU instr | LDS instr | "
           "waves/CU |", "|---|---|---|---|---|---|---|---|---|"]
    def pick(table, k):
        """The entry of kernel k, or of its one template instance (k<...>)."""
        if k in table:
            return table[k]
        inst = [v for name, v in table.items() if name.startswith(k + "<")]
        return inst[0] if len(inst) == 1 else None

    for k in kernels:
        a, b, us = pick(t1, k), pick(t2, k), pick(avg, k)
        if a is None or b is None or us is None:
            continue
        cyc = us * 1e-6 * ghz * 1e9
        cap = cus * 4 * cyc / 4
        valu = a["SQ_INSTS_VALU"]
        f64 = a["SQ_INSTS_VALU_ADD_F64"] + a["SQ_INSTS_VALU_MUL_F64"] + a["SQ_INSTS_VALU_FMA_F64"]
        waves = b["SQ_WAVE_CYCLES"] * 4 / cyc / cus
        out.append(f"| {k[5:]} | {us:.1f} | {valu:.3g} | {f64 / valu:.0%} | {valu / cap:.0%} | {f64 / cap:.0%} | "
                   f"{a['SQ_INSTS_SALU']:.3g} | {a['SQ_INSTS_LDS']:.3g} | {waves:.1f} |")
    out += ["", "Per-dispatch counter values:", ""]
    for path in (pass1, pass2):
        txt = open(path).read().split("per dispatch (sum / dispatches):")[1]
        listed = {k.rsplit("_w6", 1)[0] for k in kernels}     # both builds of the tile group kernels
        out += [x for x in txt.strip().split("\n")
                if x.startswith("|") and ("kernel" in x or "---" in x or any(k in x for k in listed))]
        out.append("")
    return "\n".join(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="?", default="")
    ap.add_argument("--counters", nargs=3, metavar=("PASS1_MD", "PASS2_MD", "TRACE_MD"),
                    help="derived SQ-counter table from two counter summaries and a kernel-trace summary")
    ap.add_argument("--counter-commands", default="",
                    help="counters mode: the rocprofv3 commands of the two passes, separated by ';;'")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--dispatches", default="", help="also list every dispatch of kernels matching this")
    ap.add_argument("--traffic-write-db", default="",
                    help="traffic mode: db is the FETCH_SIZE run, this the WRITE_SIZE run; prints JSON")
    ap.add_argument("--copy-bytes", type=int, default=256 << 20)
    ap.add_argument("--kernels", default="k_mid,k_group_fwd,k_group_inv,k_group_fwd_w6,k_group_inv_w6,k_col_fwd_s,"
                                          "k_col_fwd_b,k_col_inv_s,k_col_inv_b,k_unpack,k_carry_fix,k_combine,"
                                          "k_barrett_fix,k_basecase")
    a = ap.parse_args()
    if a.counters:
        ks = ["nbh::k_mid", "nbh::k_group_fwd_w6", "nbh::k_group_inv_w6", "nbh::k_col_fwd_s", "nbh::k_col_fwd_b",
              "nbh::k_col_inv_s", "nbh::k_col_inv_b", "nbh::k_unpack"]
        print(counters_report(*a.counters, [c for c in a.counter_commands.split(";;") if c], ks))
        return
    if not a.db:
        ap.error("a results database is needed")
    if a.traffic_write_db:
        import json
        res = traffic(a.db, a.traffic_write_db, a.kernels.split(","), a.copy_bytes)
        res["command"] = a.command
        print(json.dumps(res, indent=1))
        return
    cur = sqlite3.connect(a.db).cursor()
    t = tables(cur)
    lines = [f"# {a.title}", ""]
    if a.command:
        lines += ["Command:", "", "```", a.command, "```", ""]
    if "counters_collection" in t and cur.execute("select count(*) from counters_collection").fetchone()[0]:
        lines += counter_stats(cur)
    else:
        lines += kernel_stats(cur)
    if a.dispatches:
        lines += ["", f"dispatches of {a.dispatches}:", "",
                  "| kernel | grid | workgroup | lds | duration us |", "|---|---|---|---|---|"]
        for name, gx, wx, lds, d in cur.execute("select name, grid_x, workgroup_x, lds_size, duration from kernels "
                                                "order by start"):
            if a.dispatches in name:
                lines.append(f"| {short(name)} | {gx} | {wx} | {lds} | {d / 1e3:.1f} |")
    print("\n".join(lines))


if __name__ == "__main__":
    main()
