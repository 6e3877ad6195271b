"""Parity on the two larger workloads, over the first Newton iterations of
their step (a complete step of either runs for minutes, and past its first
iterations the Newton path depends on the linear solver's rounding):

- the 64x16x16 twist-and-fold bar (18,785 vertices): the first 12 Newton
  iterations on MI355X against the oracle's, both at PCG rtol 1e-12: same
  line-search trial count and constraint count, positions within 1e-6 of the
  bounding-box diagonal, intersection free;
- the 256 uneven cubes (n^3 cells, n from 4 to 16; 468k vertices, 2.2M tets)
  batched in one context for 20 Newton iterations: five instances (the
  largest n = 16 cube among them) equal their solo solves bit for bit, and
  three of them equal the oracle's within 1e-6 of their diagonal.
"""
import numpy as np
import pytest

from idp_amd import driver, scenes
from oracle import pyoracle
from tests import exact

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_twist_fold_first_newton_iterations_match_oracle():
    K = 12
    sc = scenes.config2()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    d2, k = pyoracle.initialize_oipc(sc.dhat, m.mass)
    say(f"[twist-fold] oracle, {K} Newton iterations")
    it_o, Xo, vo, ko, st_o = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, d2, k, max_newton=K)
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    say(f"[twist-fold] gpu, {K} Newton iterations")
    it, st = d.advance_one_time_step(max_newton=K)
    X, _ = d.state()
    d.close()
    err = np.abs(X - Xo).max() / sc.diag
    say(f"[twist-fold] newton {it} (oracle {it_o}), trials {st['line_search_trials']} "
        f"(oracle {st_o['line_search_trials']}), constraints {st['n_constraints']} (oracle {st_o['n_constraints']}), "
        f"err/diag {err:.2e}")
    assert it == it_o == K
    assert st["line_search_trials"] == st_o["line_search_trials"]
    assert st["n_constraints"] == st_o["n_constraints"]
    assert err < 1e-6
    assert not exact.surface_self_intersects(X, m.bedge, m.btri)


def test_cubes256_batch_sample_alone_and_oracle():
    K = 20
    scs = scenes.config3()
    ns = [int(round(sc.tets.shape[0] / 6) ** (1 / 3) + 0.5) for sc in scs]
    big = int(np.argmax(ns))
    small = int(np.argmin(ns))
    sample = sorted({big, small, 17, 101, 200})
    say(f"[cubes256] batch of {len(scs)}, {sum(sc.X.shape[0] for sc in scs)} vertices; sample {sample} "
        f"(n = {[ns[i] for i in sample]})")
    b = driver.BatchAnimFix(scs, pcg_rtol=1e-12)
    its, st = b.advance_one_time_step(max_newton=K)
    assert int(its.max()) <= K
    res = {i: b.state(i) for i in sample}
    b.close()
    for i in sample:
        a = driver.BatchAnimFix([scs[i]], pcg_rtol=1e-12)
        ia, _ = a.advance_one_time_step(max_newton=K)
        Xa, va = a.state(0)
        a.close()
        say(f"[cubes256] instance {i} (n={ns[i]}): newton {its[i]} alone {ia[0]}")
        assert its[i] == ia[0]
        assert np.array_equal(res[i][0], Xa) and np.array_equal(res[i][1], va)
    for i in (big, small, 101):
        sc = scs[i]
        m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        d2, k = pyoracle.initialize_oipc(sc.dhat, m.mass)
        it_o, Xo, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, d2, k, max_newton=K)
        err = np.abs(res[i][0] - Xo).max() / sc.diag
        say(f"[cubes256] instance {i} (n={ns[i]}): oracle newton {it_o}, err/diag {err:.2e}")
        assert its[i] == it_o
        assert err < 1e-6
        assert not exact.surface_self_intersects(res[i][0], m.bedge, m.btri)
