"""Continuous batching on MI355X: requests started in free slots while other
slots are mid-solve must give, bit for bit, the result of solving them alone."""
import numpy as np
import pytest

from idp_amd import capi, driver, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _slots():
    return [scenes.small_bend(), scenes.cube_sinusoid(4, seed=3), scenes.cube_sinusoid(5, seed=7),
            scenes.small_bend(angle_deg=200.0)]


def test_serve_staggered_starts_equal_alone():
    scs = _slots()
    b = driver.BatchAnimFix(scs)
    ctx = b.ctx
    ctx.serve_configure(b.h, pcg_rtol=b.pcg_rtol)
    k3 = b.kappa.copy()

    def start(ids):
        say(f"[serve] start slots {ids}")
        ctx.serve_start(ids, [scs[i].X for i in ids], [scs[i].velocity for i in ids], b.tol[ids], b.dHat2[ids],
                        k3[ids])

    start([0, 1])
    done = {}
    it = 0
    pending = [2, 3]
    while True:
        n, fin = ctx.serve_iterate()
        it += 1
        for i in np.nonzero(fin)[0]:
            done[int(i)] = ctx.serve_result(int(i))
        if it == 3 and pending:  # join while slots 0 and 1 are mid-solve
            start(pending)
            pending = []
        if n == 0 and not pending:
            break
    assert sorted(done) == [0, 1, 2, 3]
    for i, sc in enumerate(scs):
        a = driver.BatchAnimFix([sc])
        ia, st = a.advance_one_time_step()
        Xa, va = a.state(0)
        X, v, kap, sst = done[i]
        say(f"[serve] slot {i} {sc.name}: newton {sst['newton_iters']} alone {ia[0]}")
        assert sst["newton_iters"] == ia[0]
        assert np.array_equal(X, Xa) and np.array_equal(v, va)
        assert kap[0] == a.kappa[0, 0]
