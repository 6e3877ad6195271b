"""Lagged friction on MI355X against the oracle (Library/FEM/FRICTION.h and
the friction loop of Shell/IMPLICIT_EULER.h:413-437, 757-848): the basis
update (friction set bit for bit, closest points, bases and normal forces
within 1e-12), the energy, gradient and local Hessians, and whole steps of a
block on a tilted slab that slides or sticks (same Newton counts, friction
set sizes and basis updates, positions within 1e-6 of the bounding-box
diagonal)."""
import numpy as np
import pytest

from idp_amd import capi, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _slab_state(theta=20.0, mu=0.1, steps=2):
    """The block after a few oracle steps with friction (in contact, moving)."""
    sc, ns, dbc = scenes.block_on_slab()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    th = np.radians(theta)
    b = m.mass[:, None] * (9.81 * np.array([np.sin(th), -np.cos(th), 0.0]))[None, :]
    X, v, k = sc.X.copy(), np.zeros_like(sc.X), kappa.copy()
    for _ in range(steps):
        _, X, v, k, _ = m.advance_friction(X, v, sc.h, sc.newton_tol, dHat2, k, mu, 1e-6, fric_iters=0,
                                           dbc_idx=dbc, body_force=b)
    return sc, ns, dbc, m, dHat2, k, b, X, v


def _ctx(sc, dbc=None):
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    ctx = capi.Context(0)
    ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
    if dbc is not None:
        ctx.set_dbc(dbc)
    return ctx


def test_friction_stage_terms_match_oracle():
    sc, ns, dbc, m, dHat2, k, b, Xn, v = _slab_state()
    cs = m.constraint_set(Xn, dHat2)
    fs = pyoracle.FrictionSet(Xn, cs, dHat2, k[0])
    assert fs.n > 0
    ctx = _ctx(sc)
    try:
        ctx.set_friction(0.3, 1e-6, 1)
        say(f"[friction] basis of {cs.shape[0]} constraints on the device")
        nf = ctx.friction_basis(Xn, Xn, cs, dHat2, k[0])
        assert nf == fs.n
        fcs, cp, basis, N = ctx.get_friction(nf)
        assert np.array_equal(fcs.ravel(), fs.fcs)
        assert np.abs(cp.ravel() - fs.cp).max() <= 1e-12 * max(1.0, np.abs(fs.cp).max())
        assert np.abs(basis.ravel() - fs.basis).max() <= 1e-13
        assert np.abs(N - fs.nforce).max() <= 1e-12 * np.abs(fs.nforce).max()
        h = sc.h
        epsvh2 = 1e-6 * h * h
        # a displacement with sliding and static contacts
        X = Xn + np.random.default_rng(0).normal(0, 2e-5, Xn.shape) + np.array([3e-5, 0.0, 0.0])
        X[dbc] = Xn[dbc]
        say("[friction] energy, gradient and Hessian on the device")
        E = ctx.friction_energy(X, h)
        Eo = fs.energy(X, Xn, epsvh2, 0.3)
        assert E == pytest.approx(Eo, rel=1e-12)
        g = ctx.friction_gradient(X, h)
        go = fs.gradient(X, Xn, epsvh2, 0.3)
        assert np.abs(g - go).max() <= 1e-12 * np.abs(go).max()
        H = ctx.friction_hessian_local(X, h, nf)
        Ho, _ = fs.hessian_local(X, Xn, epsvh2, 0.3)
        assert np.abs(H - Ho).max() <= 1e-11 * np.abs(Ho).max()
    finally:
        ctx.close()


def test_friction_coefficients_per_component_match_oracle():
    sc, ns, dbc, m, dHat2, k, b, Xn, v = _slab_state(steps=1)
    cs = m.constraint_set(Xn, dHat2)
    comp_range = np.array([ns, sc.X.shape[0]], np.int32)  # slab, block
    mu_comp = np.array([0.0, 0.2, 0.4, 0.0])
    fs = pyoracle.FrictionSet(Xn, cs, dHat2, k[0])
    fs.coef(comp_range, mu_comp)
    ctx = _ctx(sc)
    try:
        ctx.set_friction(0.0, 1e-6, 1, comp_range, mu_comp)
        nf = ctx.friction_basis(Xn, Xn, cs, dHat2, k[0])
        _, _, _, N = ctx.get_friction(nf)
        assert np.abs(N - fs.nforce).max() <= 1e-12 * np.abs(fs.nforce).max()
        with pytest.raises(capi.IdpError):  # a vertex outside every component
            ctx.set_friction(0.0, 1e-6, 1, np.array([ns], np.int32), np.array([1.0]))
            ctx.friction_basis(Xn, Xn, cs, dHat2, k[0])
    finally:
        ctx.close()


def test_newton_system_with_friction_solves_against_oracle_terms():
    """The device Newton direction with the friction set of idp_friction_basis:
    its gradient is the frictionless one plus the oracle's friction gradient."""
    sc, ns, dbc, m, dHat2, k, b, Xn, v = _slab_state()
    X = Xn + np.array([2e-5, 0.0, 0.0])
    X[dbc] = Xn[dbc]
    Xt = (Xn + sc.h * v + sc.h ** 2 * b / m.mass[:, None]).ravel()
    r2 = dHat2
    ctx = _ctx(sc, dbc)
    try:
        g0, _, _, _ = ctx.newton_direction(X, Xt, sc.h, r2, k[0], pcg_rtol=1e-12)
        ctx.set_friction(0.3, 1e-6, 1)
        cs = ctx.constraint_set(Xn, r2)
        ctx.friction_basis(Xn, Xn, cs, r2, k[0])
        g1, sol, rel, _ = ctx.newton_direction(X, Xt, sc.h, r2, k[0], pcg_rtol=1e-12)
        fs = pyoracle.FrictionSet(Xn, cs, r2, k[0])
        gf = fs.gradient(X, Xn, 1e-6 * sc.h * sc.h, 0.3).reshape(-1, 3)
        gf[dbc] = 0.0
        tol = 1e-9 * np.abs(gf).max() + 1e-13 * np.abs(g0).max()
        assert np.abs((g1 - g0) - gf.ravel()).max() <= tol
        assert rel <= 1e-12
    finally:
        ctx.close()


@pytest.mark.parametrize("mu,fric_iters", [(0.1, 0), (0.5, 1), (0.5, 3)])
def test_friction_steps_match_oracle(mu, fric_iters):
    sc, ns, dbc = scenes.block_on_slab()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    th = np.radians(20.0)
    b = m.mass[:, None] * (9.81 * np.array([np.sin(th), -np.cos(th), 0.0]))[None, :]
    ctx = _ctx(sc, dbc)
    dHat2, kappa = capi.initialize_oipc(sc.dhat, m.mass)
    Xo, vo, ko = sc.X.copy(), np.zeros_like(sc.X), kappa.copy()
    kg = kappa.copy()
    try:
        ctx.set_friction(mu, 1e-6, fric_iters)
        ctx.set_state(sc.X, np.zeros_like(sc.X))
        for s in range(4):
            it_o, Xo, vo, ko, st_o = m.advance_friction(Xo, vo, sc.h, sc.newton_tol, dHat2, ko, mu, 1e-6,
                                                        fric_iters=fric_iters, dbc_idx=dbc, body_force=b)
            say(f"[friction mu {mu} iters {fric_iters}] step {s}: gpu")
            it, st = ctx.advance(sc.h, sc.newton_tol, dHat2, kg, 0.0, True, 1e-12, body_force=b)
            X, v = ctx.get_state()
            err = np.abs(X - Xo).max() / sc.diag
            say(f"[friction mu {mu} iters {fric_iters}] step {s}: newton {it} (oracle {it_o}), friction set "
                f"{st['n_friction']} (oracle {st_o['n_friction']}), updates {st['friction_updates']} "
                f"(oracle {st_o['friction_updates']}), trials {st['line_search_trials']} "
                f"(oracle {st_o['line_search_trials']}), err/diag {err:.2e}")
            assert it == it_o
            assert st["n_friction"] == st_o["n_friction"]
            assert st["friction_updates"] == st_o["friction_updates"]
            assert kg[0] == ko[0]
            assert err < 1e-6
            Xo, vo = X.copy(), v.copy()  # continue both from the device state
        assert st["n_friction"] > 0
        if mu == 0.1:  # sliding: the block keeps speeding up along the slope
            assert v[ns:, 0].mean() > 0.05
    finally:
        ctx.close()


def test_friction_is_refused_by_batched_solves():
    sc, ns, dbc = scenes.block_on_slab()
    ctx = _ctx(sc)
    try:
        ctx.set_friction(0.2, 1e-6, 1)
        with pytest.raises(capi.IdpError):
            ctx.serve_configure(sc.h)
    finally:
        ctx.close()
