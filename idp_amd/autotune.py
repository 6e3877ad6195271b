"""Autotuner for the launch configuration on this device.

The PCG runs two kernels per iteration.  Two of its launch parameters are
tuned:
- the SpMV shape (index into the library's table: G = 4, 8, 16, 32 lanes per
  block row, with the scalar reduction in wavefront 0 for 0..3 or in a fifth
  wavefront for 4..7);
- the update kernel's 64-row blocks per workgroup (1, 2, 4, 8).

The two kernels are timed separately (kernel-accurate HIP events on sampled
iterations), so the search is separable: every SpMV shape at the default
update width, then every update width at the best shape.  The objective is
the kernel time per PCG iteration, summed over the workloads given, because
the iteration count itself moves slightly with G (G sets the order of each
block row's sum).  For the same reason the result is one configuration for
the whole process, not one per mesh size: an instance must get the same
arithmetic alone and in any batch.

    python -m idp_amd.autotune --out idp_amd/tuning.json [--workloads bar_twist_fold,cubes32]

writes the choice and every measurement as JSON; ``capi.Context`` applies it
to each new context (``IDP_TUNING`` names another file).

The barrier Hessian's projection stage takes one constraint per team of 8,
16, 32 or 64 lanes; its time per launch is measured for each width at the
chosen PCG configuration and the fastest (summed over the workloads) is
written as ``barrier_team``.  No sum depends on the width, so the results do
not either.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import capi, driver, scenes

SHAPE_TABLE = [(4, False), (8, False), (16, False), (32, False), (4, True), (8, True), (16, True), (32, True)]
UPD_WAVES = [1, 2, 4, 8]
DEFAULT = (5, 4)
TEAMS = [3, 8, 16, 32, 64]
DEFAULT_TEAM = 3


def _workload(name):
    if name == "bar_twist_fold":
        return "one", scenes.config2()
    if name == "small_bend":
        return "one", scenes.small_bend()
    if name.startswith("cubes"):
        n = int(name[5:] or 32)
        return "batch", scenes.config3()[:n]
    raise ValueError(f"unknown workload {name}")


def measure(kind, sc, spmv_shape, upd_waves, newton):
    """Kernel time per launch (us) of the SpMV and update kernels over the
    first `newton` Newton iterations of one step."""
    d = driver.TetAnimFix(sc) if kind == "one" else driver.BatchAnimFix(sc)
    d.ctx.set_launch_config(spmv_shape, upd_waves)
    d.ctx.profile(True)
    its, _ = d.advance_one_time_step(max_newton=newton)
    p = d.ctx.profile_read()
    d.ctx.profile(False)
    sp, up = p["pcg_spmv"], p["pcg_update"]
    return {
        "spmv_shape": spmv_shape,
        "upd_waves": upd_waves,
        "spmv_us": 1e3 * sp["ms"] / max(sp["launches"], 1),
        "update_us": 1e3 * up["ms"] / max(up["launches"], 1),
        "samples": int(sp["launches"]),
    }


def measure_team(kind, sc, team, newton, config):
    """Barrier-Hessian kernel time per launch (us, both stages and the kind
    grouping) over the first `newton` Newton iterations, at `team` lanes per
    constraint and the PCG configuration `config`."""
    d = driver.TetAnimFix(sc) if kind == "one" else driver.BatchAnimFix(sc)
    d.ctx.set_launch_config(*config)
    d.ctx.set_barrier_team(team)
    d.ctx.profile(True)
    d.advance_one_time_step(max_newton=newton)
    p = d.ctx.profile_read()
    d.ctx.profile(False)
    d.close()
    bh = p["barrier_hess"]
    return {"barrier_team": team, "barrier_hess_us": 1e3 * bh["ms"] / max(bh["launches"], 1),
            "launches": int(bh["launches"])}


def choose_team(rows):
    acc = {}
    for r in rows:
        acc[r["barrier_team"]] = acc.get(r["barrier_team"], 0.0) + r["barrier_hess_us"]
    if not acc:
        return DEFAULT_TEAM
    return min(acc, key=lambda k: (acc[k], k != DEFAULT_TEAM))


def tune_team(workloads, newton, config, log=print):
    rows = []
    for name in workloads:
        kind, sc = _workload(name)
        for t in TEAMS:
            r = measure_team(kind, sc, t, newton, config)
            r["workload"] = name
            rows.append(r)
            log(f"[tune] {name} barrier team {t}: {r['barrier_hess_us']:.1f} us per launch")
    return choose_team(rows), rows


def choose(rows):
    """Best (spmv_shape, upd_waves) from measurement rows of one or more
    workloads: the shape with the least SpMV time summed over workloads
    (rows at the default update width), then the update width with the
    least update time at that shape."""
    def total(key, pred):
        acc = {}
        for r in rows:
            if pred(r):
                k = r["spmv_shape"] if key == "spmv_us" else r["upd_waves"]
                acc[k] = acc.get(k, 0.0) + r[key]
        return acc

    sh = total("spmv_us", lambda r: r["upd_waves"] == DEFAULT[1])
    if not sh:
        return DEFAULT
    best_shape = min(sh, key=lambda k: (sh[k], k != DEFAULT[0]))
    uw = total("update_us", lambda r: r["spmv_shape"] == best_shape)
    best_w = min(uw, key=lambda k: (uw[k], k != DEFAULT[1])) if uw else DEFAULT[1]
    return best_shape, best_w


def tune(workloads, newton, log=print):
    rows = []
    for name in workloads:
        kind, sc = _workload(name)
        for shape in range(len(SHAPE_TABLE)):
            r = measure(kind, sc, shape, DEFAULT[1], newton)
            r["workload"] = name
            rows.append(r)
            log(f"[tune] {name} shape {shape} {SHAPE_TABLE[shape]}: spmv {r['spmv_us']:.2f} us")
    best_shape, _ = choose(rows)
    for name in workloads:
        kind, sc = _workload(name)
        for w in UPD_WAVES:
            if w == DEFAULT[1]:
                continue
            r = measure(kind, sc, best_shape, w, newton)
            r["workload"] = name
            rows.append(r)
            log(f"[tune] {name} update width {w}: update {r['update_us']:.2f} us")
    return choose(rows), rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", default=capi.TUNING_PATH)
    ap.add_argument("--workloads", default="bar_twist_fold,cubes32")
    ap.add_argument("--newton", type=int, default=20, help="Newton iterations timed per configuration")
    a = ap.parse_args(argv)
    names = [w for w in a.workloads.split(",") if w]
    t0 = time.time()
    lg = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    (shape, waves), rows = tune(names, a.newton, log=lg)
    team, team_rows = tune_team(names, a.newton, (shape, waves), log=lg)
    out = {
        "arch": "gfx950",
        "spmv_shape": shape,
        "spmv_lanes_per_row": SHAPE_TABLE[shape][0],
        "spmv_scalar_wavefront": SHAPE_TABLE[shape][1],
        "upd_waves": waves,
        "barrier_team": team,
        "workloads": names,
        "newton_per_run": a.newton,
        "seconds": round(time.time() - t0, 1),
        "measurements": rows,
        "barrier_team_measurements": team_rows,
    }
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("spmv_shape", "spmv_lanes_per_row", "spmv_scalar_wavefront", "upd_waves",
                                          "barrier_team")}))


if __name__ == "__main__":
    main()
