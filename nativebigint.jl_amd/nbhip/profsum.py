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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", default="")
    ap.add_argument("--dispatches", default="", help="also list every dispatch of kernels matching this")
    a = ap.parse_args()
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
