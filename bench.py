#!/usr/bin/env python3
"""Throughput of the Newton-with-barrier step on one MI355X.

One step is one complete ``Advance_One_Step_IE_Discrete_Shell`` solve
(Library/FEM/Shell/IMPLICIT_EULER.h:152-891) of the animation-fix workload:
the 64x16x16-cell Kuhn bar (98,304 tets, 18,785 vertices, seeded rest jitter)
driven towards two twists plus a 180-degree fold, FP64, dhat = 1e-3 and
Newton tolerance 1e-4 of the bounding-box diagonal.  Every timed solve starts
from the same device-resident state (restored device-to-device), so the
inputs are in HBM when the clock starts.

Printed: one JSON line with solves/s, p50/p99 solve latency, the roofline
position of the dominant kernel (device time from HIP events recorded on the
library's stream around each launch, algorithmic bytes or flops from the
kernel's work model) and the oracle's time on this host's cores for a bounded
sample of the same workload.

``--max-newton K`` truncates each solve to K Newton iterations: used only for
profiler runs (the JSON line then names the truncation).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

T_START = time.monotonic()
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK = 8.0e12        # B/s, MI355X HBM3E
FP64_VEC_PEAK = 78.6e12  # FLOP/s, MI355X FP64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    from idp_amd import scenes

    if name == "bar_twist_fold":
        return scenes.config2()
    if name == "bar_bend":
        return scenes.config1()
    if name == "bar_crumple":
        return scenes.config5()
    raise SystemExit(f"unknown workload {name}")


def read_traffic(path, kernel):
    """HBM bytes per dispatch of `kernel` from a committed counter summary
    (a line "traffic_bytes_per_dispatch <kernel>: <bytes>"), or None."""
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_dispatch\s+" + re.escape(kernel) + r":\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def _nproc():
    """`nproc` of this process (it honours OMP_NUM_THREADS and the CPU affinity)."""
    import subprocess

    try:
        return max(1, int(subprocess.run(["nproc"], capture_output=True, text=True, check=True).stdout.strip()))
    except (OSError, ValueError, subprocess.CalledProcessError):
        return max(1, len(os.sched_getaffinity(0)))


def _cpu_model():
    import subprocess

    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.CalledProcessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_run(sc, threads, max_newton, rtol):
    """One oracle solve (or its first max_newton Newton iterations) on `threads`
    OpenMP threads, in a child process so that the thread count takes effect;
    returns (Newton iterations, seconds)."""
    import subprocess

    code = (
        "import sys, time, json\n"
        f"sys.path.insert(0, {HERE!r})\n"
        "from idp_amd import scenes\n"
        "from oracle import pyoracle\n"
        f"sc = scenes.{sc}()\n"
        "m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)\n"
        f"pyoracle.set_accd_max_steps({ACCD_CAP})\n"
        "d2, k = pyoracle.initialize_oipc(sc.dhat, m.mass)\n"
        "t0 = time.perf_counter()\n"
        f"it = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, d2, k, lin_rtol={rtol!r}, max_newton={max_newton})[0]\n"
        "print(json.dumps([it, time.perf_counter() - t0]))\n"
    )
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, check=True).stdout
    it, t = json.loads(out.strip().splitlines()[-1])
    return int(it), float(t)


def cpu_baseline(gpu_newton, seconds, rtol, cfg1_gpu_s):
    """The oracle (the C++/OpenMP restatement of the reference's step) on this
    host's cores, beside the GPU run:
      - the bench workload (64x16x16 twist-and-fold bar): the first M Newton
        iterations (M doubled until the sample takes about `seconds`/2), on
        `nproc` threads and on 1 thread, scaled to one solve by the GPU
        solve's Newton count (early iterations carry fewer contacts, so the
        estimate favours the CPU);
      - the 20x4x4 bend bar (the reference's CPU-runnable case): complete
        solves on `nproc` threads and on 1 thread, measured, not scaled."""
    nproc = _nproc()
    res = {}
    for thr in (nproc, 1):
        M, it, t = 1, 0, 0.0
        while True:
            it, t = _oracle_run("config2", thr, M, rtol)
            if t >= seconds / 2 / 4 or it < M or M >= 64:
                break
            M *= 2
        res[thr] = (it, t)
        log(f"[bench] cpu baseline: twist-and-fold bar, {thr} thread(s), first {it} Newton iterations in {t:.2f} s")
    full = {}
    for thr in (nproc, 1):
        it, t = _oracle_run("config1", thr, 0, rtol)
        full[thr] = {"newton": it, "seconds": t}
        log(f"[bench] cpu baseline: bend bar complete solve, {thr} thread(s): {it} Newton iterations in {t:.2f} s")
    it, t = res[nproc]
    per_iter = t / max(it, 1)
    it1, t1 = res[1]
    return {
        "value": 1.0 / (per_iter * gpu_newton),
        "unit": "solves/s",
        "cores": nproc,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": (f"oracle (C++/OpenMP restatement of the reference step; block-Jacobi PCG to the same relative "
                   f"residual in place of CHOLMOD) on the twist-and-fold bar: first {it} Newton iterations in "
                   f"{t:.2f} s on {nproc} threads ({per_iter:.3f} s/iteration), scaled by the GPU solve's "
                   f"{gpu_newton} Newton iterations"),
        "one_thread": {"value": 1.0 / (t1 / max(it1, 1) * gpu_newton), "newton_sampled": it1, "seconds": t1},
        "bend_bar_full_solve": {"threads": nproc, "cpu": full[nproc], "cpu_1_thread": full[1],
                                "gpu_seconds": cfg1_gpu_s},
    }


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB (the runtime's blit kernel):
    the known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    rc = hip.hipMemcpy(b, a, C.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    t = time.perf_counter() - t0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "rc": rc, "seconds": t}), flush=True)


def bench_batch(a):
    """256 independent cube solves batched in one context vs the plain
    one-mesh-at-a-time path (timed on an evenly spaced sample of instances)."""
    from idp_amd import capi, driver, scenes

    scs = scenes.config3()
    log(f"[bench] cubes256: {sum(x.X.shape[0] for x in scs)} vertices, {sum(x.tets.shape[0] for x in scs)} tets")
    b = driver.BatchAnimFix(scs, pcg_rtol=a.pcg_rtol)
    k0 = b.kappa.copy()
    b.ctx.snapshot()

    def solve(max_newton):
        b.ctx.restore()
        b.kappa[:] = k0
        t0 = time.perf_counter()
        its, st = b.advance_one_time_step(max_newton=max_newton, verbose=10)
        return its, st, time.perf_counter() - t0

    solve(2)  # buffer priming
    for w in range(a.warmup):
        solve(a.max_newton)
    lat, newton = [], []
    b.ctx.profile(True)
    t_start = time.perf_counter()
    for k in range(a.steps):
        its, st, t = solve(a.max_newton)
        lat.append(t)
        newton.append(int(its.max()))
        log(f"[bench] batch step {k}: {len(scs)} solves in {t:.2f} s, Newton max {its.max()} mean {its.mean():.1f}, "
            f"PCG max {max(x['pcg_iters'] for x in st)}")
    total = time.perf_counter() - t_start
    prof = b.ctx.profile_read()
    b.ctx.profile(False)
    kern = {k: {"ms": round(prof[k]["ms"], 3), "launches": prof[k]["launches"]} for k in capi.K_FAMILIES}
    for k in ("pcg_spmv", "pcg_update", "assembly"):
        if prof[k]["ms"] > 0:
            kern[k]["gbs"] = prof[k]["work"] / (prof[k]["ms"] * 1e-3) / 1e9
    if prof["barrier_hess"]["ms"] > 0:
        kern["barrier_hess"]["tflops"] = prof["barrier_hess"]["work"] / (prof["barrier_hess"]["ms"] * 1e-3) / 1e12
    kern["ccd"]["steps"] = prof["ccd_steps"]
    kern["ccd"]["max_steps"] = prof["ccd_max_steps"]
    batched = a.steps * len(scs) / total
    # plain path: the same instances one at a time
    idx = list(range(0, len(scs), max(1, len(scs) // a.plain_sample)))[: a.plain_sample]
    tp = 0.0
    for i in idx:
        d = driver.TetAnimFix(scs[i], pcg_rtol=a.pcg_rtol)
        d.ctx.snapshot()
        d.advance_one_time_step(max_newton=2)  # same priming as the batch
        d.ctx.restore()
        d.kappa[:] = k0[i]
        t0 = time.perf_counter()
        it, _ = d.advance_one_time_step(max_newton=a.max_newton)
        tp += time.perf_counter() - t0
        log(f"[bench] plain instance {i}: {it} Newton iterations, {tp:.2f} s so far")
    plain = len(idx) / tp
    log(f"[bench] plain path: {len(idx)} instances in {tp:.2f} s")
    # the batch cannot end before its slowest instance: time that one alone too
    slow = int(np.argmax(its))
    d = driver.TetAnimFix(scs[slow], pcg_rtol=a.pcg_rtol)
    d.ctx.snapshot()
    d.advance_one_time_step(max_newton=2)
    d.ctx.restore()
    d.kappa[:] = k0[slow]
    t0 = time.perf_counter()
    it_slow, st_slow = d.advance_one_time_step(max_newton=a.max_newton, verbose=200)
    t_slow = time.perf_counter() - t0
    log(f"[bench] slowest instance {slow} ({scs[slow].name}) alone: {it_slow} Newton iterations, "
        f"{st_slow['pcg_iters']} PCG iterations, {t_slow:.2f} s")
    top = np.argsort(-np.asarray(its))[:5]
    per_inst = [{"instance": int(i), "scene": scs[i].name, "vertices": int(scs[i].X.shape[0]),
                 "newton": int(its[i]), "pcg": int(st[i]["pcg_iters"])} for i in top]
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "solves/sec + p50/p99 latency at fixed accuracy; barrier pair-evals/sec vs plain",
        "value": batched,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": total * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cubes256_batched", "instances": len(scs), "pcg_rtol": a.pcg_rtol,
                   "newton_max": newton},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
        "plain_path": {"solves_per_s": plain, "sample": f"{len(idx)} evenly spaced instances, one at a time"},
        "batched_vs_plain": batched / plain,
        "slowest_alone": {"instance": slow, "newton": int(it_slow), "pcg": int(st_slow["pcg_iters"]),
                          "seconds": t_slow},
        "slowest_in_batch": per_inst,
        "kernels": kern,
    }
    print(json.dumps(out), flush=True)


def serve_sequences(frames, total=200, seed=4, count=12):
    """Animation-fix sequences: `count` bars whose size is drawn (seeded) from
    {8x4x4, 32x8x8, 64x16x16} cells, each with a target that bends it from 0
    to 270 degrees over a `total`-frame sequence (bend radius 0.15 of the bar
    length); the first `frames` frames of each are requested.  Returns
    (kinds {name: rest scene}, [(kind, [target of frame 1..frames])])."""
    from idp_amd import scenes

    sizes = ((8, 4, 4), (32, 8, 8), (64, 16, 16))
    rng = np.random.default_rng(seed)
    draw = rng.integers(0, len(sizes), count)
    draw[: len(sizes)] = np.arange(len(sizes))  # every size at least once
    kinds, seqs, cache = {}, [], {}
    for d in draw:
        nx, ny, nz = sizes[d]
        name = f"bar{nx}x{ny}x{nz}"
        if name not in cache:
            L, W = nx * 0.05, ny * 0.05
            mid = np.array([L / 2, W / 2, 0.0])  # the bend map is centred on the bar's middle
            kinds[name] = scenes.bar_bend(nx, ny, nz, 0.05, 0.15 * L, 0.0)
            cache[name] = [scenes.bar_bend(nx, ny, nz, 0.05, 0.15 * L, 270.0 * k / total).target + mid
                           for k in range(1, frames + 1)]
        seqs.append((name, cache[name]))
    return kinds, seqs


def bench_serve(a):
    """Continuous batching of animation-fix frames arriving as a Poisson
    stream, admitted by a device-memory budget (idp_amd.serve.FrameServer):
    slots per bar size are planned from the measured per-slot footprint, a
    frame waits until its predecessor in the sequence has finished and a slot
    of its size is free, and every frame warm-starts from the previous
    frame's solution.  The plain path serves the same stream one frame at a
    time (FIFO, service times measured with the one-mesh path)."""
    from idp_amd import capi, driver, serve

    kinds, seqs = serve_sequences(a.frames)
    S, F = len(seqs), a.frames
    demand = {k: sum(1 for q, _ in seqs if q == k) for k in kinds}
    budget = int(a.hbm_budget_gb * 1e9)
    srv = serve.FrameServer(kinds, budget, demand=demand, pcg_rtol=a.pcg_rtol)
    log(f"[bench] serve: {S} sequences {demand}, footprints {srv.footprint}, budget {budget} B -> slots "
        f"{srv.slots_per_kind}, resident {srv.resident_bytes} B")
    h = srv.batch.h
    k_init = {k: srv.batch.kappa[srv.slot_kind.index(k)].copy() for k in kinds}

    def run(rate, seed, frames=F):
        rng = np.random.default_rng(seed)
        order = [(s, k) for k in range(frames) for s in range(S)]
        arrive = np.cumsum(rng.exponential(1.0 / rate, len(order)))
        arr = {o: t for o, t in zip(order, arrive)}
        state = [kinds[seqs[s][0]].X.copy() for s in range(S)]
        kap = [k_init[seqs[s][0]].copy() for s in range(S)]
        nxt = [0] * S
        busy = [False] * S
        lat, waits, t0 = [], [], time.perf_counter()
        done, last = 0, 0.0
        while done < S * frames:
            now = time.perf_counter() - t0
            if now - last > 30.0:
                log(f"[bench] serve rate {rate}: {done} of {S * frames} frames after {now:.0f} s")
                last = now
            for s in range(S):
                if not busy[s] and nxt[s] < frames and arr[(s, nxt[s])] <= now:
                    q = seqs[s][0]
                    srv.submit((s, nxt[s]), q, state[s], (seqs[s][1][nxt[s]] - state[s]) / h, kap[s])
                    busy[s] = True
            if srv.pending():
                for d in srv.iterate():
                    s, k = d.tag
                    state[s] = d.X
                    kap[s][0] = d.kappa3[0]
                    lat.append(time.perf_counter() - t0 - arr[(s, k)])
                    waits.append(d.waited_iters)
                    nxt[s] += 1
                    busy[s] = False
                    done += 1
            else:
                pend = [arr[(s, nxt[s])] for s in range(S) if nxt[s] < frames]
                time.sleep(max(0.0, min(pend) - (time.perf_counter() - t0)))
        span = time.perf_counter() - t0
        return S * frames / span, np.array(lat) * 1e3, np.array(waits)

    # warm-up pass (buffers reach their working sizes): the first frame of
    # every sequence, from the rest state again afterwards
    run(1e6, 0, frames=1)
    results = {}
    for rate in [float(r) for r in a.rates.split(",")]:
        thr, lat, waits = run(rate, 1)
        results[rate] = {"solves_per_s": thr, "p50_ms": float(np.percentile(lat, 50)),
                         "p99_ms": float(np.percentile(lat, 99)), "frames_that_waited_for_a_slot":
                         int((waits > 0).sum()), "max_wait_newton_iters": int(waits.max())}
        log(f"[bench] serve rate {rate}: {thr:.2f} solves/s, p50 {results[rate]['p50_ms']:.1f} ms, "
            f"p99 {results[rate]['p99_ms']:.1f} ms, {results[rate]['frames_that_waited_for_a_slot']} frames "
            f"waited for a slot")
    # plain path: one frame at a time, FIFO over the same arrival stream
    svc = {}
    plain_ctx = {}
    for s in range(S):
        q = seqs[s][0]
        if q not in plain_ctx:
            d = driver.TetAnimFix(kinds[q], pcg_rtol=a.pcg_rtol)
            d.advance_one_time_step(max_newton=1)  # buffer priming
            plain_ctx[q] = d
        d = plain_ctx[q]
        d.kappa[:] = k_init[q]
        X = kinds[q].X
        for k in range(F):
            d.ctx.set_state(X, (seqs[s][1][k] - X) / h)
            t0 = time.perf_counter()
            d.advance_one_time_step()
            svc[(s, k)] = time.perf_counter() - t0
            X, _ = d.state()
    for d in plain_ctx.values():
        d.close()
    plain = {}
    for rate in results:
        rng = np.random.default_rng(1)
        order = [(s, k) for k in range(F) for s in range(S)]
        arrive = np.cumsum(rng.exponential(1.0 / rate, len(order)))
        t, lat = 0.0, []
        for o, ta in zip(order, arrive):
            t = max(t, ta) + svc[o]
            lat.append(t - ta)
        lat = np.array(lat) * 1e3
        plain[rate] = {"solves_per_s": len(order) / t, "p50_ms": float(np.percentile(lat, 50)),
                       "p99_ms": float(np.percentile(lat, 99))}
    top = max(results)
    out = {
        "metric": "solves/sec + p50/p99 latency at fixed accuracy; barrier pair-evals/sec vs plain",
        "value": results[top]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": S * F,
        "warmup": S,
        "ms_per_step": 1e3 / results[top]["solves_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serve_bar_frames", "sequences": S, "frames": F, "kinds": demand,
                   "slots": srv.slots_per_kind, "slot_footprint_bytes": srv.footprint, "hbm_budget_bytes": budget,
                   "resident_bytes": srv.resident_bytes, "rates": list(results), "pcg_rtol": a.pcg_rtol},
        "latency_ms": {"p50": results[top]["p50_ms"], "p99": results[top]["p99_ms"]},
        "served": results,
        "plain_path": plain,
        "served_vs_plain": results[top]["solves_per_s"] / plain[top]["solves_per_s"],
    }
    srv.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="bar_twist_fold",
                    choices=["bar_twist_fold", "bar_bend", "bar_crumple", "cubes256", "serve"],
                    help="bar_crumple (160x48x48, 2.2M tets) is a measurement workload: use it with --max-newton")
    ap.add_argument("--frames", type=int, default=200,
                    help="serve: frames requested per sequence (the first of a 200-frame bend to 270 degrees)")
    ap.add_argument("--hbm-budget-gb", type=float, default=2.0,
                    help="serve: device memory the slots may hold (slots per bar size are planned from it)")
    ap.add_argument("--rates", default="20,200", help="serve: Poisson arrival rates (req/s)")
    ap.add_argument("--plain-sample", type=int, default=16,
                    help="cubes256: instances solved one at a time for the plain-path comparison")
    ap.add_argument("--pcg-rtol", type=float, default=1e-8)
    ap.add_argument("--max-newton", type=int, default=0)
    ap.add_argument("--budget-s", type=float, default=480.0,
                    help="wall-clock budget from process start: no warm-up or timed solve starts unless it "
                         "and the closing measurements still fit (at least one solve is always timed)")
    ap.add_argument("--reserve-s", type=float, default=60.0,
                    help="time kept for the closing measurements (profiled solve, CPU baseline)")
    ap.add_argument("--profile-newton", type=int, default=60,
                    help="Newton iterations of the profiled solve that gives the per-family device times")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(HERE, "profiles", "k_pcg_spmv_traffic_bar_twist_fold.txt"))
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the sharded layout is described in the design notes, not launched", file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return

    from idp_amd import capi, driver

    if a.workload == "cubes256":
        return bench_batch(a)
    if a.workload == "serve":
        return bench_serve(a)
    return bench_plain(a)


def _roofline(fam, prof, traffic_path):
    """Roofline position of kernel family `fam` from a profile: algorithmic
    work of its sampled launches over their summed device time."""
    f = prof[fam]
    launch_s = f["ms"] * 1e-3 / max(f["launches"], 1)
    per_launch = f["work"] / max(f["launches"], 1)
    if fam in ("barrier_hess", "tet_hess"):
        # FP64 vector arithmetic (no MFMA instruction is issued): the FP64 vector peak
        roof = {"kernel": fam, "bound": "fp64-vector", "achieved": per_launch / launch_s / 1e12,
                "peak": FP64_VEC_PEAK / 1e12, "unit": "TFLOP/s", "traffic": None}
    else:
        roof = {"kernel": fam, "bound": "hbm", "achieved": per_launch / launch_s / 1e9, "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "traffic": read_traffic(traffic_path, fam)}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["avg_launch_ms"] = launch_s * 1e3
    roof["algorithmic_per_launch"] = per_launch
    roof["launches_sampled"] = f["launches"]
    return roof


# ACCD step cap of every solve of the bench (the library default), applied to
# the oracle's CPU baseline too
ACCD_CAP = 1 << 20


def bench_plain(a):
    """One step = one complete solve of the workload from the same device
    state.  Warm-up and timed solves run while the remaining wall-clock
    budget (--budget-s, counted from process start) still holds one more
    solve and the closing measurements; at least one solve is timed.  The
    timed solves record only the sampled PCG kernel times (profile level 2);
    a profiled partial solve afterwards gives the per-family breakdown."""
    import signal

    from idp_amd import capi, driver, scenes

    sc = workload(a.workload)
    log(f"[bench] workload {sc.name}: {sc.X.shape[0]} vertices, {sc.tets.shape[0]} tets, budget {a.budget_s:.0f} s")
    d = driver.TetAnimFix(sc, pcg_rtol=a.pcg_rtol)
    k0 = d.kappa.copy()
    d.ctx.snapshot()
    state = {"lat": [], "its": [], "pcg": [], "capped": [], "warm": 0, "done": False}

    def solve(max_newton):
        d.ctx.restore()
        d.kappa[:] = k0
        t0 = time.perf_counter()
        # returns after the context's stream has drained (stats and kappa are copied back)
        it, st = d.advance_one_time_step(max_newton=max_newton, verbose=100)
        return it, st, time.perf_counter() - t0

    def line(partial):
        lat, its = state["lat"], state["its"]
        n = len(lat)
        total = float(sum(lat))
        lat_ms = np.array(lat) * 1e3 if n else np.array([float("nan")])
        out = {
            "metric": "solves/sec + p50/p99 latency at fixed accuracy; barrier pair-evals/sec vs plain",
            "value": n / total if n else 0.0,
            "unit": "solves/s",
            "n_gpus": 1,
            "steps": n,
            "warmup": state["warm"],
            "ms_per_step": total * 1e3 / n if n else None,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": sc.name, "vertices": int(sc.X.shape[0]), "tets": int(sc.tets.shape[0]),
                       "pcg_rtol": a.pcg_rtol, "newton_iters": its, "pcg_iters": state["pcg"],
                       "accd_capped_pairs": state["capped"], "accd_max_steps": ACCD_CAP,
                       "max_newton": a.max_newton or None, "precond": capi.PRECOND,
                       "launch_config": list(d.ctx.launch_config()),
                       "steps_requested": a.steps, "warmup_requested": a.warmup, "budget_s": a.budget_s},
            "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99))},
            # wall time per stage of each timed solve (host clock around each stage, ms)
            "stage_ms": state.get("stages", []),
        }
        out.update(state.get("extra", {}))
        if partial:
            out["partial"] = partial
        return out

    def on_term(signum, frame):
        if not state["done"]:
            state["done"] = True
            print(json.dumps(line(f"stopped by signal {signum} before the closing measurements")), flush=True)
        sys.exit(128 + signum)

    signal.signal(signal.SIGTERM, on_term)
    elapsed = lambda: time.monotonic() - T_START  # noqa: E731
    reserve = a.reserve_s  # closing measurements: profiled partial solve, CPU baseline

    # buffer priming (device allocations grow on first use): a 2-iteration solve
    solve(2)
    est = None
    for w in range(a.warmup):
        # timed solves come first: a further warm-up runs only while up to
        # three timed solves would still fit after it
        if est is not None and elapsed() + est * (1 + min(a.steps, 3)) + reserve > a.budget_s:
            log(f"[bench] warm-up stopped after {w}: budget")
            break
        it, st, t = solve(a.max_newton)
        est = t if est is None else max(est, t)
        state["warm"] += 1
        log(f"[bench] warmup {w}: {it} Newton iterations, {st['pcg_iters']} PCG iterations, {t:.2f} s")

    d.ctx.profile(2)  # sampled PCG kernel pairs only: the launch stream runs as it would unprofiled
    for k in range(a.steps):
        if k > 0 and elapsed() + est + reserve > a.budget_s:
            log(f"[bench] timed solves stopped after {k}: budget")
            break
        it, st, t = solve(a.max_newton)
        est = t if est is None else max(est, t)
        state["lat"].append(t)
        state["its"].append(it)
        state["pcg"].append(st["pcg_iters"])
        state["capped"].append(st["accd_capped"])
        state.setdefault("stages", []).append({k2[3:]: round(st[k2], 1) for k2 in (
            "ms_constraint_set", "ms_ccd", "ms_energy", "ms_gradient", "ms_hessian", "ms_solve", "ms_total")})
        log(f"[bench] step {k}: {it} Newton iterations, {st['pcg_iters']} PCG iterations, "
            f"{st['n_constraints']} constraints, {st['accd_capped']} capped ACCD pairs, {t:.2f} s")
    prof_t = d.ctx.profile_read()
    d.ctx.profile(0)
    roof = _roofline("pcg_spmv", prof_t, a.traffic)
    extra = {"roofline": roof}
    extra["roofline_update_kernel"] = _roofline("pcg_update", prof_t, a.traffic)
    state["extra"] = extra

    # per-family device times: a profiled solve truncated to a bounded number
    # of Newton iterations (an event pair around every launch)
    nprof = a.profile_newton
    d.ctx.profile(1)
    it_p, st_p, t_p = solve(nprof)
    prof = d.ctx.profile_read()
    d.ctx.profile(0)
    kern = {}
    for k in capi.K_FAMILIES:
        e = prof[k]
        kern[k] = {"ms": round(e["ms"], 3), "launches": e["launches"]}
    bh = prof["barrier_hess"]
    if bh["ms"] > 0:
        kern["barrier_hess"]["tflops"] = bh["work"] / (bh["ms"] * 1e-3) / 1e12
        kern["barrier_hess"]["frac_fp64_peak"] = bh["work"] / (bh["ms"] * 1e-3) / FP64_VEC_PEAK
    for k in ("pcg", "pcg_spmv", "pcg_update", "assembly"):
        if prof[k]["ms"] > 0:
            kern[k]["gbs"] = prof[k]["work"] / (prof[k]["ms"] * 1e-3) / 1e9
    pairs = int(sum(prof["barrier_pairs"]))
    kern["barrier_hess"]["pairs"] = pairs
    kern["barrier_hess"]["clamped"] = int(sum(prof["barrier_clamped"]))
    # evaluations per constraint kind: PT, EE, PE, PP, mollified EE, PE, PP
    kern["barrier_hess"]["pairs_by_kind"] = [int(x) for x in prof["barrier_pairs"]]
    kern["ccd"]["steps"] = prof["ccd_steps"]
    kern["ccd"]["max_steps"] = prof["ccd_max_steps"]
    # what a graph of the PCG iterations could remove at most: the PCG stage's
    # device time between its kernels (launch gaps, chunk-boundary waits and
    # the preconditioner build), from the sampled per-launch kernel times
    sp, up = prof["pcg_spmv"], prof["pcg_update"]
    if sp["launches"] and up["launches"] and prof["pcg"]["ms"] > 0:
        per_it = sp["ms"] / sp["launches"] + up["ms"] / up["launches"]
        kms = per_it * st_p["pcg_iters"]
        kern["pcg"]["iterations"] = int(st_p["pcg_iters"])
        kern["pcg"]["kernel_ms_est"] = kms
        kern["pcg"]["outside_kernels_ms"] = prof["pcg"]["ms"] - kms
        kern["pcg"]["outside_kernels_frac"] = (prof["pcg"]["ms"] - kms) / prof["pcg"]["ms"]
    extra["kernels_profiled_solve"] = {"newton_iters": it_p, "seconds": t_p, "families": kern}
    extra["barrier_hessian_pair_evals_per_s"] = (pairs / (bh["ms"] * 1e-3)) if bh["ms"] > 0 else None
    log(f"[bench] profiled solve: {it_p} Newton iterations in {t_p:.2f} s")

    # the reference's CPU-runnable case on the GPU, complete solve (for the CPU comparison)
    d1 = driver.TetAnimFix(scenes.config1(), pcg_rtol=a.pcg_rtol)
    d1.ctx.snapshot()
    d1.advance_one_time_step()
    d1.ctx.restore()
    d1.kappa[:] = capi.initialize_oipc(d1.scene.dhat, d1.mass)[1]
    t0 = time.perf_counter()
    d1.advance_one_time_step()
    cfg1_gpu_s = time.perf_counter() - t0
    d1.close()

    if not a.no_cpu_baseline and state["its"]:
        log(f"[bench] cpu baseline (oracle, about {a.cpu_seconds:.0f} s)")
        extra["cpu_baseline"] = cpu_baseline(int(np.median(state["its"])), a.cpu_seconds, a.pcg_rtol, cfg1_gpu_s)
    else:
        extra["cpu_baseline"] = None
    state["done"] = True
    print(json.dumps(line(None)), flush=True)
    d.close()


if __name__ == "__main__":
    main()
