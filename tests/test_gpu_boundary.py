"""The step's boundary on MI355X, through the C ABI, against the oracle:
body force and the flow variant of idp_advance_one_step (the Hinge and Flow
bindings, DISCRETE_SHELL.h:1110, 1113; IMPLICIT_EULER.h:196-217, 850-854),
the ACCD step cap, the zero-step error (IPC.h:2014-2031), the d^2 <= 0 error
(IPC.h:773-776), the gradient-descent fallback after a failed linear solve
(DIRECT_SOLVER.h:42-44, IMPLICIT_EULER.h:531-551), and the argument checks of
the serving entry point.

Bars: same Newton count and end state within 1e-6 of the bounding-box
diagonal; CCD step sizes bit for bit.
"""
import numpy as np
import pytest

from idp_amd import capi, driver, scenes
from oracle import pyoracle
from tests import exact
from tests.conftest import ORACLE_ACCD_CAP

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _solve_both(sc, body_force, flow, vel=None):
    vel = sc.velocity if vel is None else vel
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    d2, k = pyoracle.initialize_oipc(sc.dhat, m.mass)
    it_o, Xo, vo, ko, st_o = m.advance(sc.X, vel, sc.h, sc.newton_tol, d2, k, body_force=body_force, flow=flow)
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    d.ctx.set_state(sc.X, vel)
    it, st = d.advance_one_time_step(body_force=body_force, flow=flow)
    X, v = d.state()
    d.close()
    return (it_o, Xo, vo, ko, st_o), (it, X, v, d.kappa, st), m


def test_gravity_hinge_step_matches_oracle():
    sc = scenes.small_bend()
    m0 = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    b = m0.mass[:, None] * np.array([0.0, -9.81, 0.0])[None, :]
    say("[boundary] hinge step with gravity")
    (it_o, Xo, vo, _, st_o), (it, X, v, _, st), m = _solve_both(sc, b, False)
    err = np.abs(X - Xo).max() / sc.diag
    say(f"[boundary] gravity: gpu newton {it} oracle {it_o}, err/diag {err:.2e}")
    assert it == it_o
    assert err < 1e-6
    assert not exact.surface_self_intersects(X, m.bedge, m.btri)
    # the force moved the result: compare with the zero-force step
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    d.advance_one_time_step()
    X0, _ = d.state()
    d.close()
    assert np.abs(X - X0).max() > 1e-3 * sc.h ** 2 * 9.81


def test_flow_step_in_contact_matches_oracle():
    """Flow variant held in contact: the body force pulls the upper of two
    stacked blocks down into the lower one and sideways (v is zeroed, the tet
    term is off); the barrier holds it, so the loop runs until the stopping
    rule with contacts left (15 Newton iterations in the oracle)."""
    sc, _ = scenes.two_blocks_slide()
    m0 = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    nA = sc.X.shape[0] // 2
    acc = np.zeros_like(sc.X)
    acc[nA:, 0] = -0.3
    acc[nA:, 1] = -1.0
    b = m0.mass[:, None] * acc
    say("[boundary] flow step pressed into contact")
    (it_o, Xo, vo, _, st_o), (it, X, v, _, st), m = _solve_both(sc, b, True, vel=np.zeros_like(sc.X))
    err = np.abs(X - Xo).max() / sc.diag
    say(f"[boundary] flow in contact: gpu newton {it} oracle {it_o}, constraints {st['n_constraints']} "
        f"(oracle {st_o['n_constraints']}), err/diag {err:.2e}")
    assert it == it_o > 1
    assert err < 1e-6
    assert st["n_constraints"] == st_o["n_constraints"] > 0
    assert not exact.surface_self_intersects(X, m.bedge, m.btri)
    assert np.abs(v - (X - sc.X) / sc.h).max() <= 1e-9 * np.abs(v).max()


def test_flow_step_ends_when_no_contact_is_left():
    """Flow variant toward the self-intersecting bend target of a small bar:
    the CCD-limited first step leaves every surface pair farther apart than
    dhat, so the loop ends after one iteration (IMPLICIT_EULER.h:850-854) in
    both implementations."""
    sc = scenes.small_bend()
    m0 = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    b = m0.mass[:, None] * (sc.target - sc.X) / sc.h ** 2
    say("[boundary] flow step toward the bend target")
    (it_o, Xo, vo, _, st_o), (it, X, v, _, st), m = _solve_both(sc, b, True)
    err = np.abs(X - Xo).max() / sc.diag
    say(f"[boundary] flow: gpu newton {it} oracle {it_o}, constraints {st['n_constraints']}, err/diag {err:.2e}")
    assert it == it_o == 1
    assert st["n_constraints"] == st_o["n_constraints"] == 0
    assert err < 1e-6
    assert not exact.surface_self_intersects(X, m.bedge, m.btri)


def test_flow_without_contact_is_one_newton_step_to_xtilde():
    X, T = scenes.kuhn_block(2, 2, 2, 0.05)
    sc = scenes.Scene("block", X.copy(), T.astype(np.int32), X.copy())
    m0 = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    acc = np.zeros_like(sc.X)
    acc[:, 0] = 3.0 * sc.X[:, 1] / sc.h ** 2
    b = m0.mass[:, None] * acc
    say("[boundary] flow step without contact")
    (it_o, Xo, _, _, _), (it, Xg, vg, _, st), _ = _solve_both(sc, b, True, vel=np.full_like(sc.X, 0.7))
    want = sc.X + sc.h ** 2 * acc
    assert it == it_o == 1
    assert st["n_constraints"] == 0
    assert np.abs(Xg - want).max() <= 1e-12 * np.abs(want).max()


def test_accd_step_cap_matches_oracle_bit_for_bit():
    sc, p = scenes.two_blocks_slide()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    ctx = capi.Context(0)
    try:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        out = {}
        # every pair stepped: a pair the no-hit certificate would end is
        # otherwise stopped by the cap and taken as a hit
        ctx.set_accd_certificate(False)
        for cap in (200, 1000, -1):
            say(f"[boundary] ACCD cap {cap}")
            ctx.set_accd_max_steps(cap)
            pyoracle.set_accd_max_steps(max(cap, 0))
            ag, _ = ctx.ccd(sc.X.ravel(), p.ravel(), 1.0)
            ao, _ = m.ccd(sc.X, p, 1.0)
            assert ag == ao
            out[cap] = ag
        assert 0 < out[200] < out[1000] < out[-1] <= 1.0
        # with the certificate (the default) and no cap: the reference's loop
        ctx.set_accd_certificate(True)
        ag, _ = ctx.ccd(sc.X.ravel(), p.ravel(), 1.0)
        assert ag == out[-1]
    finally:
        pyoracle.set_accd_max_steps(ORACLE_ACCD_CAP)
        ctx.close()


def test_capped_pairs_are_counted_in_the_step_stats():
    # gap above dhat: no contact at the start, so the first Newton direction
    # is the slide itself (two rigid translations); each facing pair then
    # needs about 1 / (0.9 * 0.01) = 111 additive steps
    sc, p = scenes.two_blocks_slide(gap=0.01)
    assert 0.01 * 0.05 > sc.dhat
    sc.target = sc.X + p
    d = driver.TetAnimFix(sc, pcg_rtol=1e-10)
    say("[boundary] capped pairs counted")
    it, st = d.ctx.advance(sc.h, sc.newton_tol, d.dHat2, d.kappa, max_newton=1, accd_max_steps=50)
    assert st["accd_capped"] > 0
    d.close()
    d = driver.TetAnimFix(sc, pcg_rtol=1e-10)
    it, st = d.ctx.advance(sc.h, sc.newton_tol, d.dHat2, d.kappa, max_newton=1, accd_max_steps=-1)
    assert st["accd_capped"] == 0
    d.close()


def test_zero_ccd_step_is_reported():
    sc = scenes.small_bend()
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    with capi.Context(0) as ctx:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        say("[boundary] zero CCD step")
        with pytest.raises(capi.IdpError) as e:
            ctx.ccd(sc.X.ravel(), (sc.target - sc.X).ravel(), 0.0)
        assert e.value.code == capi.IDP_E_CCD_ZERO


def test_nonpositive_distance_is_reported():
    """Thickness above the gap between two blocks: d^2 - t^2 <= 0 in a barrier
    term at the start of the step and in the stage gradient."""
    sc, _ = scenes.two_blocks_slide(gap=2e-4)  # 1e-5 apart
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = capi.initialize_oipc(sc.dhat, mass)
    th = 2e-5
    with capi.Context(0) as ctx:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        ctx.set_state(sc.X, np.zeros_like(sc.X))
        say("[boundary] d^2 <= 0 at the start of a step")
        with pytest.raises(capi.IdpError) as e:
            ctx.advance(sc.h, sc.newton_tol, dHat2, kappa, th)
        assert e.value.code == capi.IDP_E_NONPOS_DIST
        cs = ctx.constraint_set(sc.X, dHat2, th)
        assert cs.shape[0] > 0
        with pytest.raises(capi.IdpError) as e:
            ctx.barrier_gradient(sc.X, cs, dHat2, kappa[0], th)
        assert e.value.code == capi.IDP_E_NONPOS_DIST


def test_gradient_descent_fallback_matches_oracle():
    """A small tet far from the bar with no stiffness and a negative mass on
    its vertices makes the Newton matrix indefinite where the gradient lives:
    the linear solve fails (the device's preconditioner finds a block that is
    not positive definite, the oracle's PCG breaks down with p.q < 0) and the
    step takes the gradient direction instead (DIRECT_SOLVER.h:42-44,
    IMPLICIT_EULER.h:531-551).  Each such iteration doubles that tet's offset
    from x~; the bar, at rest at x~, stays put."""
    sc = scenes.small_bend()
    nb = sc.X.shape[0]
    tiny = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.0], [0.0, 0.0, 0.01]]) + 1.0
    X = np.vstack([sc.X, tiny])
    T = np.vstack([sc.tets, np.arange(nb, nb + 4)[None, :]]).astype(np.int32)
    vol, IB, lam, mu, mass = capi.fem_init(X, T, sc.density, sc.E, sc.nu)
    lam[-1] = mu[-1] = 0.0
    mass[nb:] = -1.0
    v = np.zeros_like(X)
    v[nb:] = [1.0, 0.0, 0.0]
    m = pyoracle.Mesh(X, T, sc.density, sc.E, sc.nu)
    m.lam[-1] = m.mu[-1] = 0.0
    m.mass[nb:] = -1.0
    dHat2, kappa = capi.initialize_oipc(sc.dhat, mass[:nb])
    it_o, Xo, _, _, _ = m.advance(X, v, sc.h, sc.newton_tol, dHat2, kappa.copy(), max_newton=3)
    cold = capi.PCG_COLD
    capi.PCG_COLD = 1  # every solve from zero, as the oracle's
    try:
        with capi.Context(0) as ctx:
            ctx.set_mesh(X, T, IB, vol, lam, mu, mass)
            ctx.set_state(X, v)
            say("[boundary] gradient-descent fallback")
            it, st = ctx.advance(sc.h, sc.newton_tol, dHat2, kappa.copy(), pcg_rtol=1e-12, max_newton=3)
            Xg, _ = ctx.get_state()
    finally:
        capi.PCG_COLD = cold
    assert it == it_o == 3
    assert st["gd_iters"] == 3
    assert np.abs(Xg - Xo).max() <= 1e-12
    # x - x~ = -h v doubles each iteration: -h, -2h, -4h, -8h
    assert np.allclose(Xg[nb:, 0], X[nb:, 0] - 7 * sc.h, rtol=0, atol=1e-12)
    assert np.abs(Xg[:nb] - X[:nb]).max() <= 1e-15  # the rest shape's rounding-level gradient


def test_serve_start_rejects_duplicate_and_busy_slots():
    scs = [scenes.small_bend(), scenes.small_bend(angle_deg=200.0)]
    b = driver.BatchAnimFix(scs)
    ctx = b.ctx
    ctx.serve_configure(b.h)
    X = [sc.X for sc in scs]
    V = [sc.velocity for sc in scs]
    say("[boundary] serve_start argument checks")
    with pytest.raises(capi.IdpError):
        ctx.serve_start([0, 0], [X[0], X[0]], [V[0], V[0]], b.tol[[0, 0]], b.dHat2[[0, 0]], b.kappa[[0, 0]])
    ctx.serve_start([0], [X[0]], [V[0]], b.tol[[0]], b.dHat2[[0]], b.kappa[[0]])
    with pytest.raises(capi.IdpError):
        ctx.serve_start([0], [X[0]], [V[0]], b.tol[[0]], b.dHat2[[0]], b.kappa[[0]])  # busy
    n = 1
    for _ in range(2000):
        n, fin = ctx.serve_iterate()
        if n == 0:
            break
    assert n == 0
    b.close()
