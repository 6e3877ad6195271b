"""Batched independent solves on MI355X, through the C ABI.

An instance's result must not depend on the batch it is solved in: every
reduction (energies, dot products, PCG scalars, CCD minima, grid set-up) is
taken over the instance alone, in the order a one-instance solve uses, so the
final positions are compared bit for bit.  A batch of one must reproduce the
plain one-mesh step exactly, and each instance must match the oracle's solve
to 1e-6 of its bounding-box diagonal.
"""
import numpy as np
import pytest

from idp_amd import driver, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _mix():
    return [scenes.small_bend(), scenes.cube_sinusoid(4, seed=3), scenes.config1(), scenes.cube_sinusoid(5, seed=7)]


def test_batch_of_one_equals_plain():
    sc = scenes.small_bend()
    say("[batch] plain vs batch of one")
    d = driver.TetAnimFix(sc)
    it, _ = d.advance_one_time_step()
    X1, v1 = d.state()
    b = driver.BatchAnimFix([sc])
    its, st = b.advance_one_time_step()
    X2, v2 = b.state(0)
    assert its[0] == it
    assert np.array_equal(X1, X2) and np.array_equal(v1, v2)
    assert b.kappa[0, 0] == d.kappa[0]


def test_batch_alone_equals_mix():
    mix = _mix()
    say(f"[batch] mix of {len(mix)}")
    b = driver.BatchAnimFix(mix)
    its, st = b.advance_one_time_step()
    say(f"[batch] newton {list(its)}")
    for i, sc in enumerate(mix):
        say(f"[batch] alone {sc.name}")
        a = driver.BatchAnimFix([sc])
        ia, _ = a.advance_one_time_step()
        Xa, va = a.state(0)
        Xm, vm = b.state(i)
        assert ia[0] == its[i]
        assert np.array_equal(Xa, Xm) and np.array_equal(va, vm)
        assert a.kappa[0, 0] == b.kappa[i, 0]


def test_batch_matches_oracle():
    mix = [scenes.cube_sinusoid(4, seed=3), scenes.cube_sinusoid(5, seed=7)]
    b = driver.BatchAnimFix(mix, pcg_rtol=1e-12)
    its, _ = b.advance_one_time_step()
    for i, sc in enumerate(mix):
        m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
        it, Xo, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa)
        X, _ = b.state(i)
        err = np.abs(X - Xo).max() / sc.diag
        say(f"[batch] {sc.name}: gpu newton {its[i]} oracle {it} err/diag {err:.2e}")
        assert err < 1e-6


def test_batch_handover_changes_nothing():
    """A tiny ACCD step budget hands long pairs from one Newton iteration to
    the next, and PCG solves still running when half of the instances' solves
    have stopped are continued in the next one (the instance waits while the
    others go on); the results must be those of a batch that never hands
    over, bit for bit."""
    mix = _mix()
    say("[batch] hand-over: ACCD budget 4 and PCG hand-over vs never")
    a = driver.BatchAnimFix(mix)
    ia, _ = a.advance_one_time_step(accd_budget=4, pcg_handover=0)
    b = driver.BatchAnimFix(mix)
    ib, _ = b.advance_one_time_step(accd_budget=-1, pcg_handover=-1)
    say(f"[batch] newton {list(ia)} vs {list(ib)}")
    assert list(ia) == list(ib)
    for i in range(len(mix)):
        Xa, va = a.state(i)
        Xb, vb = b.state(i)
        assert np.array_equal(Xa, Xb) and np.array_equal(va, vb)
        assert a.kappa[i, 0] == b.kappa[i, 0]


def test_256_cube_batch_instances_equal_their_solo_solves():
    """The 256-cube batch of the batched workload (cubes of 4^3..16^3 cells,
    seeded sinusoid targets), first 20 Newton iterations of every instance:
    four sampled instances (the smallest, the largest and two between)
    solved alone give the same iterates bit for bit."""
    scs = scenes.config3()
    assert len(scs) == 256
    say("[batch] 256 cubes, 20 Newton iterations")
    b = driver.BatchAnimFix(scs)
    its, _ = b.advance_one_time_step(max_newton=20)
    assert min(its) >= 1 and max(its) <= 20
    sizes = np.array([sc.X.shape[0] for sc in scs])
    pick = [int(np.argmin(sizes)), int(np.argmax(sizes)), 37, 200]
    for i in pick:
        say(f"[batch] instance {i} ({scs[i].name}, {sizes[i]} vertices) alone")
        a = driver.BatchAnimFix([scs[i]])
        ia, _ = a.advance_one_time_step(max_newton=20)
        Xa, va = a.state(0)
        Xm, vm = b.state(i)
        assert ia[0] == its[i]
        assert np.array_equal(Xa, Xm) and np.array_equal(va, vm)
        a.close()
    b.close()
