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


def test_serve_admission_waits_for_memory_then_equals_alone():
    """A device-memory budget that holds two slots of the bar's kind; three
    sequences ask for a frame at once.  The third waits until a slot is free,
    runs in a slot another sequence used before it, and every result equals,
    bit for bit, the frame solved alone."""
    from idp_amd import serve

    seqs = [scenes.small_bend(angle_deg=a) for a in (200.0, 230.0, 250.0)]
    rest = seqs[0]
    fp = serve.slot_footprint(rest)
    say(f"[serve] one slot of {rest.name}: {fp} device bytes")
    assert fp > 0
    srv = serve.FrameServer({"bar": rest}, budget_bytes=int(2.5 * fp), demand={"bar": 3})
    try:
        assert srv.slots_per_kind == {"bar": 2}
        k0 = srv.batch.kappa[0]
        for i, sc in enumerate(seqs):
            srv.submit(i, "bar", sc.X, sc.velocity, k0)
        done = {}
        while srv.pending():
            for d in srv.iterate():
                done[d.tag] = d
        say(f"[serve] waited (Newton iterations): {[done[i].waited_iters for i in range(3)]}, "
            f"slots {[done[i].slot for i in range(3)]}, resident {srv.resident_bytes} B, budget {srv.budget} B")
        assert sorted(done) == [0, 1, 2]
        assert done[0].waited_iters == 0 and done[1].waited_iters == 0
        assert done[2].waited_iters > 0
        assert done[2].slot in (done[0].slot, done[1].slot)
        assert len(srv.slot_users[done[2].slot]) == 2
        for i, sc in enumerate(seqs):
            a = driver.BatchAnimFix([sc])
            ia, st = a.advance_one_time_step()
            Xa, va = a.state(0)
            a.close()
            d = done[i]
            say(f"[serve] sequence {i}: newton {d.stats['newton_iters']} alone {ia[0]}")
            assert d.stats["newton_iters"] == ia[0]
            assert np.array_equal(d.X, Xa) and np.array_equal(d.v, va)
    finally:
        srv.close()
