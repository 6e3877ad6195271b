"""Launch-gap breakdown of the PCG solves from a rocprofv3 kernel trace.

    rocprofv3 --kernel-trace --output-format csv -d <dir> -o gap -- \\
        python3 -m idp_amd.trace_gaps solve --max-newton 150
    python3 -m idp_amd.trace_gaps analyze <dir>/gap_kernel_trace.csv

`solve` runs the first Newton iterations of the 64x16x16 twist-and-fold bar
(one mesh, the default solver options).  `analyze` groups the trace's
consecutive PCG kernels (SpMV, update, their partial-sum reductions and the coarse level) into
runs, one run per chunk of enqueued iterations, and reports the kernels'
device time against the runs' span: the difference is the time between
kernels inside a run, which is what capturing the chunk in a graph could
remove at most.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import sys

import numpy as np

PCG_KERNELS = ("k_pcg_spmv", "k_pcg_update", "k_pcg_red_spmv", "k_pcg_red_upd", "k_coarse_restrict", "k_coarse_solve")


def _base(name):
    n = name.split("(")[0].replace("void ", "").replace("idp::", "")
    return n.split("<")[0]


def analyze(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    tot, cnt = collections.Counter(), collections.Counter()
    for r in rows:
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        tot[_base(r["Kernel_Name"])] += d
        cnt[_base(r["Kernel_Name"])] += 1
    runs, cur = [], []
    for r in rows:
        if _base(r["Kernel_Name"]) in PCG_KERNELS:
            cur.append(r)
        elif cur:
            runs.append(cur)
            cur = []
    if cur:
        runs.append(cur)
    span = busy = 0
    gaps = []
    for run in runs:
        span += int(run[-1]["End_Timestamp"]) - int(run[0]["Start_Timestamp"])
        busy += sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in run)
        gaps += [int(b["Start_Timestamp"]) - int(a["End_Timestamp"]) for a, b in zip(run, run[1:])]
    g = np.array(gaps, dtype=np.float64) / 1e3 if gaps else np.zeros(1)
    # PCG launches after a solve's stopping iteration return at once: the
    # working ones are those above 1.5 x the shortest
    working = {}
    for k in ("k_pcg_spmv", "k_pcg_update", "k_coarse_restrict", "k_coarse_solve"):
        d = np.array([(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in rows
                      if _base(r["Kernel_Name"]) == k])
        if d.size:
            w = d[d >= 1.5 * d.min()]
            working[k] = {"launches": int(d.size), "working": int(w.size), "avg_us_working": float(w.mean()) if w.size else 0.0,
                          "avg_us_all": float(d.mean())}
    return {
        "trace_span_ms": (int(rows[-1]["End_Timestamp"]) - int(rows[0]["Start_Timestamp"])) / 1e6,
        "kernel_time_ms": sum(tot.values()) / 1e3,
        "top_kernels_ms_calls_us": [(k, round(v / 1e3, 2), cnt[k], round(v / cnt[k], 2))
                                    for k, v in tot.most_common(12)],
        "pcg_runs": len(runs), "pcg_span_ms": span / 1e6, "pcg_kernel_ms": busy / 1e6,
        "pcg_gap_frac": 1 - busy / span if span else 0.0,
        "gap_us_median": float(np.median(g)), "gap_us_p90": float(np.percentile(g, 90)),
        "gap_us_p99": float(np.percentile(g, 99)), "gap_us_max": float(g.max()), "gaps": len(gaps),
        "pcg_working_launches": working,
    }


def solve(max_newton):
    from . import driver, scenes

    sc = scenes.config2()
    d = driver.TetAnimFix(sc)
    it, st = d.ctx.advance(sc.h, sc.newton_tol, d.dHat2, d.kappa, max_newton=max_newton)
    print(json.dumps({"newton": it, "pcg_iters": st["pcg_iters"], "ms_total": st["ms_total"],
                      "ms_solve": st["ms_solve"], "ms_ccd": st["ms_ccd"]}), flush=True)
    d.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("solve")
    s.add_argument("--max-newton", type=int, default=150)
    a_ = sub.add_parser("analyze")
    a_.add_argument("trace")
    a = ap.parse_args(argv)
    if a.cmd == "solve":
        solve(a.max_newton)
    else:
        print(json.dumps(analyze(a.trace), indent=1))


if __name__ == "__main__":
    sys.exit(main())
