"""Kernel sweeps on one GPU: the barrier Hessian (form + projection stages)
over growing constraint counts, timed by the library's own event profiler
on its stream, with the same algorithmic flop count per constraint kind as
the solve's profile.

    python -m idp_amd.kbench barrier_hess [--max 1048576] [--reps 3]

The constraints are the canonical set of the 160x48x48 crumple bar at 90% of
its first CCD step with a contact radius of 10 dhat (about 1M constraints of
every kind); a count N takes the first N of them.  Prints one JSON line per
count.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi, scenes

FP64_VEC_PEAK = 78.6e12


def _crumple_state():
    sc = scenes.config5()
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    ctx = capi.Context(0)
    ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
    dHat2, kappa = capi.initialize_oipc(sc.dhat, mass)
    p = (sc.target - sc.X).ravel()
    alpha, _ = ctx.ccd(sc.X, p, 1.0)
    X1 = sc.X.ravel() + 0.9 * alpha * p
    r2 = 100.0 * dHat2
    cs = ctx.constraint_set(X1, r2)
    return ctx, X1, cs, r2, kappa[0]


def barrier_hess_sweep(max_n=1 << 20, reps=3, team=None):
    ctx, X, cs, r2, kappa = _crumple_state()
    if team:
        ctx.set_barrier_team(team)
    print(f"[kbench] {cs.shape[0]} constraints available, team {ctx.barrier_team()}", file=sys.stderr, flush=True)
    n = 1 << 12
    sizes = []
    while n < min(max_n, cs.shape[0]):
        sizes.append(n)
        n <<= 2
    sizes.append(min(max_n, cs.shape[0]))
    out = []
    ctx.barrier_hessian_local(X, cs[: sizes[0]], r2, kappa)  # buffers grow once
    for n in sizes:
        ctx.profile(1)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.barrier_hessian_local(X, cs[:n], r2, kappa)
        wall = time.perf_counter() - t0
        p = ctx.profile_read()
        ctx.profile(0)
        bh = p["barrier_hess"]
        ms = bh["ms"] / max(bh["launches"], 1)
        fl = bh["work"] / max(bh["launches"], 1)
        rec = {"kernel": "barrier_hess", "constraints": int(n), "launches": bh["launches"], "ms_per_launch": ms,
               "flops_per_launch": fl, "tflops": fl / (ms * 1e-3) / 1e12, "frac_fp64_peak": fl / (ms * 1e-3) /
               FP64_VEC_PEAK, "ns_per_constraint": ms * 1e6 / n, "pairs_by_kind": p["barrier_pairs"],
               "clamped_by_kind": p["barrier_clamped"], "wall_s_with_copies": wall, "team": ctx.barrier_team()}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    ctx.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel", choices=["barrier_hess"])
    ap.add_argument("--max", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--team", type=int, default=None)
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        sys.exit("kbench: no HIP device")
    barrier_hess_sweep(a.max, a.reps, a.team)


if __name__ == "__main__":
    main()
