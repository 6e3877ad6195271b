"""Kernel sweeps on one GPU: the barrier Hessian (form + projection stages)
and the friction kernels (basis update, Hessian) over growing constraint
counts, timed by the library's own event profiler on its stream, with the
same algorithmic work per constraint kind as the solve's profile.

    python -m idp_amd.kbench barrier_hess [--max 1048576] [--reps 3]
    python -m idp_amd.kbench friction [--max 1048576] [--reps 3]

The constraints are the canonical set of the 160x48x48 crumple bar at 90% of
its first CCD step with a contact radius of 10 dhat (about 1M constraints of
every kind); a count N takes the first N of them.  Prints one JSON line per
count.  The friction sweep takes the bar's rest positions as the step's
start x_n, so every contact has slid (the sliding branch of the Hessian).
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi, scenes

FP64_VEC_PEAK = 78.6e12
HBM_PEAK = 8.0e12


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
    return ctx, X1, cs, r2, kappa[0], sc


def _sizes(max_n, avail):
    n = 1 << 12
    sizes = []
    while n < min(max_n, avail):
        sizes.append(n)
        n <<= 2
    sizes.append(min(max_n, avail))
    return sizes


def friction_sweep(max_n=1 << 20, reps=3):
    """Friction basis update and Hessian kernels: algorithmic bytes per
    friction constraint of nv vertices = 96 (int4 entry + 10-double record)
    + 48 nv (x and x_n of its vertices) + 72 nv^2 (its 3 x 3 blocks) for the
    Hessian as a Newton system forms it; 24 per barrier constraint (entry,
    flag, position) + 96 + 96 per friction constraint (the vertex reads of
    its geometry, the record written) for the basis update."""
    ctx, X, cs, r2, kappa, sc = _crumple_state()
    ctx.set_friction(0.3, 1e-6, 1)
    print(f"[kbench] {cs.shape[0]} constraints available", file=sys.stderr, flush=True)
    out = []
    for n in _sizes(max_n, cs.shape[0]):
        ctx.profile(1)
        for _ in range(reps):
            nf = ctx.friction_basis(X, sc.X, cs[:n], r2, kappa)
        # the Hessian as a solve forms it (assembly blocks), inside a Newton
        # system at X with that friction set (one PCG iteration suffices)
        for _ in range(reps):
            ctx.newton_direction(X, X, sc.h, r2, kappa, pcg_rtol=0.5, pcg_maxit=1)
        p = ctx.profile_read()
        ctx.profile(0)
        fcs, _, _, _ = ctx.get_friction(nf)
        nv = np.where(fcs[:, 0] >= 0, 4, np.where(fcs[:, 3] >= 0, 4, np.where(fcs[:, 2] >= 0, 3, 2)))
        hb = float((96 + 48 * nv + 72 * nv * nv).sum())
        bb = 24.0 * n + 192.0 * nf
        fh, fb = p["friction"], p["friction_basis"]
        ms_h = fh["ms"] / max(fh["launches"], 1)
        ms_b = fb["ms"] / max(fb["launches"], 1)
        rec = {"kernel": "friction", "constraints": int(n), "friction_constraints": int(nf),
               "by_nv": {int(k): int((nv == k).sum()) for k in (2, 3, 4)},
               "hess_ms_per_launch": ms_h, "hess_bytes_per_launch": hb,
               "hess_gbs": hb / (ms_h * 1e-3) / 1e9, "hess_frac_hbm": hb / (ms_h * 1e-3) / HBM_PEAK,
               "basis_ms_per_launch": ms_b, "basis_bytes_per_launch": bb,
               "basis_gbs": bb / (ms_b * 1e-3) / 1e9, "basis_frac_hbm": bb / (ms_b * 1e-3) / HBM_PEAK}
        print(json.dumps(rec), flush=True)
        out.append(rec)
    ctx.close()
    return out


def barrier_hess_sweep(max_n=1 << 20, reps=3, team=None):
    ctx, X, cs, r2, kappa, _ = _crumple_state()
    if team:
        ctx.set_barrier_team(team)
    print(f"[kbench] {cs.shape[0]} constraints available, team {ctx.barrier_team()}", file=sys.stderr, flush=True)
    sizes = _sizes(max_n, cs.shape[0])
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
    ap.add_argument("kernel", choices=["barrier_hess", "friction"])
    ap.add_argument("--max", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--team", type=int, default=None)
    a = ap.parse_args(argv)
    if capi.device_count() < 1:
        sys.exit("kbench: no HIP device")
    if a.kernel == "friction":
        friction_sweep(a.max, a.reps)
    else:
        barrier_hess_sweep(a.max, a.reps, a.team)


if __name__ == "__main__":
    main()
