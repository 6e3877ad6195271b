"""Summaries of a rocprofv3 kernel trace (``--kernel-trace --output-format
csv``): per-kernel dispatch counts and device times, and the idle time of
the device between consecutive dispatches of chosen kernels (launch gaps).

    python -m idp_amd.profsum <run>_kernel_trace.csv [--gaps k_pcg_spmv,k_pcg_update] [--top 25]

A gap is the time from the end of one dispatch to the start of the next one
on the same queue, attributed to the kernel that starts after it; gaps longer
than --idle-ms are host waits (synchronisation, host work), not launch
overhead, and are reported apart.
"""
from __future__ import annotations

import argparse
import csv
import sys
from collections import defaultdict

import numpy as np


def load(path):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((r["Kernel_Name"], int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                         r.get("Queue_Id", "0")))
    rows.sort(key=lambda t: t[1])
    return rows


def short(name, width=60):
    n = name.split("(")[0]
    return n if len(n) <= width else n[: width - 3] + "..."


def kernel_table(rows, top):
    tot = defaultdict(list)
    for n, s, e, _ in rows:
        tot[n].append((e - s) * 1e-3)
    allt = sum(sum(v) for v in tot.values())
    out = [f"{'kernel':<60} {'calls':>9} {'total ms':>11} {'avg us':>10} {'%':>6}"]
    for n, v in sorted(tot.items(), key=lambda kv: -sum(kv[1]))[:top]:
        out.append(f"{short(n):<60} {len(v):>9} {sum(v) / 1e3:>11.2f} {np.mean(v):>10.2f} "
                   f"{100 * sum(v) / allt:>6.2f}")
    return out


def gap_table(rows, names, idle_ms):
    byq = defaultdict(list)
    for r in rows:
        byq[r[3]].append(r)
    gaps = defaultdict(list)
    busy = defaultdict(float)
    for q, rs in byq.items():
        for prev, cur in zip(rs, rs[1:]):
            if not any(k in cur[0] for k in names):
                continue
            key = next(k for k in names if k in cur[0])
            gaps[key].append((cur[1] - prev[2]) * 1e-3)  # us
            busy[key] += (cur[2] - cur[1]) * 1e-3
    out = [f"{'kernel':<24} {'dispatches':>10} {'kernel us':>10} {'gap p50 us':>11} {'gap mean us':>12} "
           f"{'gap sum ms':>11} {'host waits':>11} {'waits ms':>9}"]
    for k in names:
        g = np.array(gaps.get(k, []))
        if g.size == 0:
            continue
        launch = g[g <= idle_ms * 1e3]
        wait = g[g > idle_ms * 1e3]
        out.append(f"{k:<24} {g.size:>10} {busy[k] / g.size:>10.2f} {np.median(launch):>11.2f} "
                   f"{launch.mean():>12.2f} {launch.sum() / 1e3:>11.2f} {wait.size:>11} {wait.sum() / 1e3:>9.2f}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("trace")
    ap.add_argument("--gaps", default="k_pcg_spmv,k_pcg_update")
    ap.add_argument("--idle-ms", type=float, default=0.2)
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args(argv)
    rows = load(a.trace)
    span = (rows[-1][2] - rows[0][1]) * 1e-6 if rows else 0.0
    busy = sum(e - s for _, s, e, _ in rows) * 1e-6
    print(f"dispatches {len(rows)}, first start to last end {span:.1f} ms, kernel time {busy:.1f} ms "
          f"({100 * busy / max(span, 1e-9):.1f}% of the span)")
    print("\n".join(kernel_table(rows, a.top)))
    print()
    print(f"gaps before each dispatch (end of the previous dispatch on the queue to this start; "
          f"> {a.idle_ms} ms counted as host waits):")
    print("\n".join(gap_table(rows, a.gaps.split(","), a.idle_ms)))


if __name__ == "__main__":
    sys.exit(main())
