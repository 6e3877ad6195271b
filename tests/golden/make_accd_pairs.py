"""Long additive-CCD pairs of the 64x16x16 twist-and-fold bar, as inputs.

The device solver logs every ACCD pair that ran 65,536 steps or more when the
environment variable IDP_ACCD_LOG is set: its start positions and (mean-free)
motions, the bound it started with, its result (1 hit, 2 no hit, 3 stopped by
the step cap) and its step count.  This script keeps a few of each kind from
such a log as a fixture; the tests rebuild every pair as two triangles and
compare the step size of the device CCD (with and without the no-hit
certificate) against the oracle's, so the logged results serve only to pick a
spread of cases.

Run:
    IDP_ACCD_LOG=1 python3 -m idp_amd.trace_gaps solve --max-newton 900 2> log.txt
    python3 tests/golden/make_accd_pairs.py log.txt
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PAT = re.compile(r"long ACCD pair \((-?\d+),(-?\d+)\) inst (\d+) d0 (\S+) maxDisp (\S+) tmax (\S+) toc (\S+) "
                 r"steps (\S+) result (\S+) d (\S+)")


def main(path):
    pairs, last = [], None
    for line in open(path):
        m = PAT.search(line)
        if m:
            last = m.groups()
            continue
        if "long ACCD pair start" in line and last is not None:
            v = [float(x) for x in line.split("start")[1].split()]
            pairs.append({"ee": int(last[0]) < 0, "d0": float(last[3]), "alpha": float(last[5]),
                          "steps": int(float(last[7])), "result": int(float(last[8])),
                          "x": v[:12], "p": v[12:24]})
            last = None
    keep, seen = [], set()
    # a spread of step counts per result, each geometry once
    for res, lo, hi, n in ((2, 100_000, 400_000, 5), (1, 100_000, 400_000, 3), (3, 0, 1 << 30, 1)):
        cand = sorted((p for p in pairs if p["result"] == res and lo <= p["steps"] <= hi), key=lambda p: p["steps"])
        step = max(1, len(cand) // n)
        for p in cand[::step]:
            key = tuple(round(x, 12) for x in p["x"])
            if key in seen or sum(q["result"] == res for q in keep) >= n:
                continue
            seen.add(key)
            keep.append(p)
    out = os.path.join(HERE, "accd_long_pairs.json")
    with open(out, "w") as f:
        json.dump({"source": "IDP_ACCD_LOG of the 64x16x16 twist-and-fold bar, 825 Newton iterations",
                   "pairs": keep}, f, indent=0)
    print(f"wrote {len(keep)} pairs to {out}")


if __name__ == "__main__":
    main(sys.argv[1])
