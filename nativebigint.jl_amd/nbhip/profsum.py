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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--dispatches", default="", help="also list every dispatch of kernels matching this")
    ap.add_argument("--traffic-write-db", default="",
                    help="traffic mode: db is the FETCH_SIZE run, this the WRITE_SIZE run; prints JSON")
    ap.add_argument("--copy-bytes", type=int, default=256 << 20)
    ap.add_argument("--kernels", default="k_mid,k_group_fwd,k_group_inv,k_group_fwd_w6,k_group_inv_w6,k_unpack,"
                                          "k_carry_fix,k_combine,k_barrett_fix")
    a = ap.parse_args()
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
