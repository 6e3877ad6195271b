"""Shell normal flow on MI355X against the oracle: the surface primitives of
a shell mesh, the normal-flow force (bit for bit), the flow variant's Newton
system (gradient and BSR matrix within 1e-10), and whole flow steps of
inflating spheres pressing on each other (same Newton and line-search
counts, positions within 1e-6 of the bounding-box diagonal)."""
import numpy as np
import pytest

from idp_amd import capi, driver, scenes
from oracle import pyoracle
from tests import exact

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _ctx(sc):
    kept, mass, vol, _, _ = capi.shell_init(sc.X, sc.tris, sc.density, 0.0, 0.0, sc.thickness)
    ctx = capi.Context(0)
    ctx.set_mesh_shell(sc.X, mass, kept, vol)
    return ctx


def test_shell_surface_and_normal_flow_force():
    sc = scenes.sphere_pair()
    m = pyoracle.ShellMesh(sc.X, sc.tris, sc.density, sc.thickness)
    ctx = _ctx(sc)
    try:
        bn, be, bt = ctx.get_surface()
        assert np.array_equal(bn, m.bnode) and np.array_equal(be, m.bedge) and np.array_equal(bt, m.btri)
        Y = sc.X + np.random.default_rng(0).normal(0, 0.01, sc.X.shape)
        say("[shell] normal-flow force on the device")
        b = ctx.normal_flow_force(0.2, Y)
        assert np.array_equal(b, m.normal_flow_force(Y, 0.2))
    finally:
        ctx.close()


def test_non_flow_step_on_a_shell_without_elasticity_is_refused():
    sc = scenes.sphere_pair()
    ctx = _ctx(sc)
    try:
        ctx.set_state(sc.X, np.zeros_like(sc.X))
        k = np.array([1e5, 1e7, 0.0])
        with pytest.raises(capi.IdpError):
            ctx.advance(sc.h, sc.newton_tol, 1e-6, k, flow=False)
    finally:
        ctx.close()


def _pressed(sc, steps):
    m = pyoracle.ShellMesh(sc.X, sc.tris, sc.density, sc.thickness)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    X = sc.X.copy()
    v = np.zeros_like(X)
    for _ in range(steps):
        b = m.normal_flow_force(X, sc.magnitude)
        _, X, v, kappa, _ = m.advance(X, v, sc.h, sc.newton_tol, dHat2, kappa, body_force=b)
    return m, dHat2, kappa, X


def test_flow_newton_system_matches_oracle():
    import scipy.sparse as sps

    sc = scenes.sphere_pair()
    m, dHat2, kappa, X = _pressed(sc, 2)
    X = X.ravel()
    b = m.normal_flow_force(X, sc.magnitude)
    Xt = X + sc.h ** 2 * (b / m.mass[:, None]).ravel()
    r2 = 400 * dHat2
    rp_o, col_o, val_o, g_o = m.newton_system(X, Xt, sc.h, r2, kappa[0])
    ctx = _ctx(sc)
    try:
        say("[shell] flow Newton system on the device")
        g, sol, rel, it = ctx.newton_direction(X, Xt, sc.h, r2, kappa[0], pcg_rtol=1e-12, flow=True)
        rp, col, val = ctx.get_bsr()
    finally:
        ctx.close()
    assert np.linalg.norm(g - g_o) <= 1e-10 * np.linalg.norm(g_o)
    n = m.nV
    A = sps.bsr_matrix((val, col, rp), shape=(3 * n, 3 * n)).toarray()
    Ao = sps.bsr_matrix((val_o, col_o, rp_o), shape=(3 * n, 3 * n)).toarray()
    assert np.abs(A - Ao).max() <= 1e-10 * np.abs(Ao).max()
    assert rel <= 1e-12 and np.linalg.norm(A @ sol + g) <= 2e-12 * np.linalg.norm(g)


@pytest.mark.parametrize("counts,steps", [((2, 1, 1), 4), ((2, 2, 2), 3)])
def test_flow_steps_match_oracle(counts, steps):
    level = 2 if counts == (2, 1, 1) else 3
    sc = scenes.sphere_lattice(counts, level)
    m = pyoracle.ShellMesh(sc.X, sc.tris, sc.density, sc.thickness)
    d = driver.ShellNormalFlow(sc, pcg_rtol=1e-12)
    # the oracle starts from the driver's kappa (initialize_OIPC's mean mass is
    # summed in order there, pairwise by numpy in the oracle's helper)
    dHat2, kappa = d.dHat2, d.kappa.copy()
    assert kappa[0] == pytest.approx(pyoracle.initialize_oipc(sc.dhat, m.mass)[1][0], rel=1e-12)
    Xo = sc.X.copy()
    vo = np.zeros_like(Xo)
    try:
        for s in range(steps):
            b = m.normal_flow_force(Xo, sc.magnitude)
            it_o, Xo, vo, kappa, st_o = m.advance(Xo, vo, sc.h, sc.newton_tol, dHat2, kappa, body_force=b)
            say(f"[shell {sc.name}] step {s}: gpu")
            it, st = d.advance_one_time_step()
            X, v = d.state()
            err = np.abs(X - Xo).max() / sc.diag
            say(f"[shell {sc.name}] step {s}: newton {it} (oracle {it_o}), trials {st['line_search_trials']} "
                f"(oracle {st_o['line_search_trials']}), constraints {st['n_constraints']} "
                f"(oracle {st_o['n_constraints']}), err/diag {err:.2e}")
            assert it == it_o
            assert st["line_search_trials"] == st_o["line_search_trials"]
            assert d.kappa[0] == kappa[0]
            assert err < 1e-6
            Xo = X.copy()  # continue both from the device state
            vo = v.copy()
            d.ctx.set_state(Xo, vo)
        assert st["n_constraints"] > 0
        assert not exact.surface_self_intersects(X, m.bedge, m.btri)
    finally:
        d.close()
