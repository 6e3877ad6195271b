"""Parity of the HIP path with the oracle, through the C ABI (MI355X).

Bars: bit-exact pair sets and CCD step sizes; energies and gradients within
1e-10 relative; local Hessians within 1e-10 of their max entry (1e-9 after
the PSD clamp); Newton direction within the PCG relative residual of a
direct sparse solve; final positions within 1e-6 of the bounding-box
diagonal; every accepted iterate intersection free.
"""
import numpy as np
import pytest

from idp_amd import capi, driver, scenes
from oracle import pyoracle
from tests import exact

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _canon(cs):
    return sorted(tuple(int(x) for x in c) for c in cs)


class Case:
    def __init__(self, sc):
        self.sc = sc
        self.m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        self.dHat2, self.kappa = pyoracle.initialize_oipc(sc.dhat, self.m.mass)
        it, Xo, vo, ko, st = self.m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, self.dHat2, self.kappa)
        self.it_o, self.Xo, self.st_o = it, Xo.ravel(), st
        vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        say(f"[{sc.name}] context + set_mesh")
        self.ctx = capi.Context(0)
        self.ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)


@pytest.fixture(scope="module", params=["small", "config1"])
def case(request):
    sc = scenes.small_bend() if request.param == "small" else scenes.config1()
    return Case(sc)


def _states(c):
    sc = c.sc
    return [("solution", c.Xo), ("halfway", sc.X.ravel() + 0.5 * (c.Xo - sc.X.ravel()))]


def test_surface_sizes(case):
    say(f"[{case.sc.name}] surface sizes")
    assert case.ctx.surface_sizes() == (case.m.bnode.size, case.m.bedge.shape[0], case.m.btri.shape[0])


@pytest.mark.parametrize("radius_mult", [1, 25])
def test_constraint_set_bit_exact(case, radius_mult):
    for tag, X in _states(case):
        d2 = case.dHat2 * radius_mult
        say(f"[{case.sc.name}] constraint set at {tag} r2x{radius_mult}")
        g = case.ctx.constraint_set(X, d2)
        o = case.m.constraint_set(X, d2)
        assert _canon(g) == _canon(o)


def test_barrier_energy_gradient_hessian(case):
    X = case.Xo
    cs = case.m.constraint_set(X, case.dHat2)
    assert len(cs) > 0
    k = case.kappa[0]
    say(f"[{case.sc.name}] barrier energy/gradient/hessian on {len(cs)} constraints")
    Eg = case.ctx.barrier_energy(X, cs, case.dHat2, k)
    Eo = case.m.barrier_energy(X, cs, case.dHat2, k)
    assert abs(Eg - Eo) <= 1e-10 * abs(Eo)
    gg = case.ctx.barrier_gradient(X, cs, case.dHat2, k)
    go = case.m.barrier_gradient(X, cs, case.dHat2, k)
    assert np.linalg.norm(gg - go) <= 1e-10 * np.linalg.norm(go)
    team0 = case.ctx.barrier_team()
    try:
        # every projection-stage layout of the order-9 kinds: lane triples
        # (K in registers) and LDS teams of 8, 16 and 64 lanes
        for team in (3, 8, 16, 64):
            case.ctx.set_barrier_team(team)
            for project, tol in ((False, 1e-10), (True, 1e-9)):
                Hg, vg = case.ctx.barrier_hessian_local(X, cs, case.dHat2, k, project=project)
                Ho, vo = case.m.barrier_hessian_local(X, cs, case.dHat2, k, project=project)
                assert np.array_equal(vg, vo)
                scale = np.abs(Ho).max(1)
                assert np.all(np.abs(Hg - Ho).max(1) <= tol * scale), (team, project)
    finally:
        case.ctx.set_barrier_team(team0)


def test_min_dist2(case):
    X = case.Xo
    cs = case.m.constraint_set(X, case.dHat2)
    say(f"[{case.sc.name}] min dist2")
    dg, mg = case.ctx.min_dist2(X, cs)
    do, mo = case.m.min_dist2(X, cs)
    assert np.array_equal(dg, do) and mg == mo


def test_elastic_terms(case):
    w = case.sc.h ** 2
    nT = case.sc.tets.shape[0]
    for tag, X in _states(case):
        say(f"[{case.sc.name}] elastic terms at {tag}")
        Eg, Eo = case.ctx.elastic_energy(X, w), case.m.elastic_energy(X, w)
        assert abs(Eg - Eo) <= 1e-10 * abs(Eo)
        gg, go = case.ctx.elastic_gradient(X, w), case.m.elastic_gradient(X, w)
        assert np.linalg.norm(gg - go) <= 1e-10 * np.linalg.norm(go)
        Hg = case.ctx.elastic_hessian_local(X, w, nT)
        Ho = case.m.elastic_hessian_local(X, w)
        assert np.all(np.abs(Hg - Ho).max(1) <= 1e-9 * np.abs(Ho).max(1))


def test_elastic_invalid_state_detected(case):
    X = case.sc.X.copy().ravel()
    t = case.sc.tets[0]
    X[3 * t[0]: 3 * t[0] + 3] = X[3 * t[1]: 3 * t[1] + 3]  # collapse a tet: det F = 0
    say(f"[{case.sc.name}] degenerate tet detection")
    assert case.ctx.elastic_energy(X, 1.0) is None
    assert case.m.elastic_energy(X, 1.0) is None


def test_ccd_bit_exact(case):
    sc = case.sc
    for tag, X, p in [("rest->target", sc.X.ravel(), (sc.target - sc.X).ravel()),
                      ("solution->target", case.Xo, sc.target.ravel() - case.Xo),
                      ("rest->half", sc.X.ravel(), 0.5 * (sc.target - sc.X).ravel())]:
        say(f"[{sc.name}] CCD {tag}")
        ag, _ = case.ctx.ccd(X, p, 1.0)
        ao, _ = case.m.ccd(X, p, 1.0)
        assert ag == ao
        assert 0 < ag <= 1.0


def test_ccd_bit_exact_concentrated_motion(case):
    """A few fast vertices and a still rest: the mean motion (span cap) stays
    small, so the movers' swept boxes cover hundreds of voxels and go through
    the workgroup-per-query pass of the broad phase."""
    sc = case.sc
    X = sc.X.ravel()
    d = (sc.target - sc.X).copy()
    movers = sc.X[:, 0] <= np.quantile(sc.X[:, 0], 0.05)
    d[~movers] = 0.0
    for scale in (1.0, 4.0):
        p = scale * d.ravel()
        say(f"[{sc.name}] CCD concentrated motion x{scale}")
        ag, _ = case.ctx.ccd(X, p, 1.0)
        ao, _ = case.m.ccd(X, p, 1.0)
        assert ag == ao
        assert 0 < ag <= 1.0


def test_newton_direction_vs_direct_solve(case):
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    sc = case.sc
    X = sc.X.ravel() + 0.5 * (case.Xo - sc.X.ravel())
    xt = sc.target.ravel()
    say(f"[{sc.name}] newton direction")
    g, sol, rel, it = case.ctx.newton_direction(X, xt, sc.h, case.dHat2, case.kappa[0], pcg_rtol=1e-12)
    assert it > 0 and rel <= 1e-12
    rp, col, val = case.ctx.get_bsr()
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    A = sps.bsr_matrix((val, col, rp), shape=(3 * n, 3 * n)).tocsc()
    assert abs(A - A.T).max() <= 1e-12 * abs(A).max()
    d = spla.spsolve(A, -g)
    assert np.linalg.norm(sol - d) <= 1e-8 * np.linalg.norm(d)
    # gradient = oracle elastic + inertia + barrier
    w = sc.h ** 2
    m = case.m
    cs = m.constraint_set(X, case.dHat2)
    go = m.elastic_gradient(X, w) + np.repeat(m.mass, 3) * (X - xt)
    if len(cs):
        go += m.barrier_gradient(X, cs, case.dHat2, case.kappa[0])
    assert np.linalg.norm(g - go) <= 1e-10 * np.linalg.norm(go)
    del rows


@pytest.mark.parametrize("precond", [1, 2])
def test_newton_direction_other_preconditioners(case, precond, monkeypatch):
    """The 3x3 block Jacobi (1) and the contact-only clusters (2) reach the
    same direction as the default clusters, against the sparse direct solve."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    monkeypatch.setattr(capi, "PRECOND", precond)
    sc = case.sc
    X = sc.X.ravel() + 0.5 * (case.Xo - sc.X.ravel())
    say(f"[{sc.name}] newton direction, preconditioner {precond}")
    g, sol, rel, it = case.ctx.newton_direction(X, sc.target.ravel(), sc.h, case.dHat2, case.kappa[0],
                                                pcg_rtol=1e-12)
    assert it > 0 and rel <= 1e-12
    rp, col, val = case.ctx.get_bsr()
    n = rp.size - 1
    A = sps.bsr_matrix((val, col, rp), shape=(3 * n, 3 * n)).tocsc()
    d = spla.spsolve(A, -g)
    assert np.linalg.norm(sol - d) <= 1e-8 * np.linalg.norm(d)


def test_assembly_matches_oracle(case):
    """BSR pattern identical to the oracle's triplet assembly; values within
    1e-10 of the largest entry (block sums in a different order)."""
    sc = case.sc
    X = sc.X.ravel() + 0.5 * (case.Xo - sc.X.ravel())
    xt = sc.target.ravel()
    say(f"[{sc.name}] assembly vs oracle")
    g, _, _, _ = case.ctx.newton_direction(X, xt, sc.h, case.dHat2, case.kappa[0], pcg_rtol=1e-6)
    rp, col, val = case.ctx.get_bsr()
    orp, ocol, oval, og = case.m.newton_system(X, xt, sc.h, case.dHat2, case.kappa[0])
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.abs(val.reshape(-1, 3, 3) - oval).max() <= 1e-10 * np.abs(oval).max()
    assert np.linalg.norm(g - og) <= 1e-10 * np.linalg.norm(og)


def test_advance_matches_oracle(case):
    sc = case.sc
    say(f"[{sc.name}] advance one step (oracle newton {case.it_o})")
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    it, st = d.advance_one_time_step()
    X, v = d.state()
    err = np.abs(X.ravel() - case.Xo).max() / sc.diag
    say(f"[{sc.name}] gpu newton {it}, max err/diag {err:.3e}, constraints {st['n_constraints']}")
    assert err < 1e-6
    assert not exact.surface_self_intersects(X, case.m.bedge, case.m.btri)
    cs = case.m.constraint_set(X.ravel(), case.dHat2)
    assert case.m.min_dist2(X.ravel(), cs)[1] > 0
    assert np.allclose(v, (X - sc.X) / sc.h, rtol=0, atol=1e-12 * np.abs(v).max())


def test_repeat_is_bitwise_deterministic():
    sc = scenes.small_bend()
    say("[determinism] two identical solves")
    outs = []
    for _ in range(2):
        d = driver.TetAnimFix(sc)
        it, st = d.advance_one_time_step()
        X, _ = d.state()
        outs.append((it, X.copy(), st["pcg_iters"]))
    assert outs[0][0] == outs[1][0] and outs[0][2] == outs[1][2]
    assert np.array_equal(outs[0][1], outs[1][1])


def test_snapshot_restore():
    sc = scenes.small_bend()
    d = driver.TetAnimFix(sc)
    d.ctx.snapshot()
    k0 = d.kappa.copy()
    say("[snapshot] solve, restore, solve")
    d.advance_one_time_step()
    X1, _ = d.state()
    d.ctx.restore()
    d.kappa[:] = k0
    X0, v0 = d.state()
    assert np.array_equal(X0, sc.X)
    d.advance_one_time_step()
    X2, _ = d.state()
    assert np.array_equal(X1, X2)


def test_advance_with_dbc_matches_oracle():
    """Static DBC (one end cross-section fixed): rows/columns projected out,
    fixed vertices bit-unchanged, same Newton path as the oracle."""
    sc = scenes.small_bend()
    fixed = np.nonzero(sc.X[:, 0] <= sc.X[:, 0].min() + 1e-12)[0].astype(np.int32)
    assert fixed.size > 0
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    it_o, Xo, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=fixed)
    say(f"[dbc] {fixed.size} fixed vertices, oracle newton {it_o}")
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    d.ctx.set_dbc(fixed)
    it, st = d.advance_one_time_step()
    X, _ = d.state()
    say(f"[dbc] gpu newton {it}")
    assert it == it_o
    assert np.array_equal(X[fixed], sc.X[fixed])
    assert np.abs(X - Xo).max() / sc.diag < 1e-6


@pytest.mark.parametrize("push,gap", [(0.1, 0.5), (0.5, 0.2)])
def test_advance_with_moving_dbc_matches_oracle(push, gap):
    """Moving DBC (IMPLICIT_EULER.h:304-401, 701-755): one block's face is
    driven toward a second block.  With push < gap the whole move fits and the
    nodes are set to their targets; with push > gap the CCD cuts the move
    short, the augmented-Lagrangian penalty pulls the nodes in, its stiffness
    doubles near convergence and is dropped once they are within 1% of the
    move.  Same CCD fraction (bit for bit), Newton count and end state as the
    oracle."""
    sc, idx, tgt = scenes.two_blocks_push(push=push, gap=gap)
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    it_o, Xo, _, _, st_o = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=idx,
                                     dbc_target=tgt)
    say(f"[mdbc] push {push} gap {gap}: oracle newton {it_o}, dbc alpha {st_o['dbc_alpha']}")
    d = driver.TetAnimFix(sc, pcg_rtol=1e-12)
    d.ctx.set_dbc_targets(idx, tgt)
    it, st = d.advance_one_time_step()
    X, _ = d.state()
    say(f"[mdbc] gpu newton {it}, dbc alpha {st['dbc_alpha']}, stiffness {st['dbc_stiff']}")
    assert st["dbc_alpha"] == st_o["dbc_alpha"]
    assert st["dbc_stiff"] == st_o["dbc_stiff"] == 0.0
    if push < gap:
        assert it == it_o
        assert np.abs(X - Xo).max() / sc.diag < 1e-6
        assert np.array_equal(X[idx], tgt)
    else:
        # The penalty path ends on discrete decisions (stiffness doubling,
        # the 99% switch-off, kappa doubling), so its end state is set only to
        # about the Newton tolerance: the oracle's own end state moves by
        # 6e-4 of the diagonal between linear tolerances 1e-8 and 1e-12 here.
        err = np.abs(X - Xo).max() / sc.diag
        say(f"[mdbc] end state vs oracle: {err:.2e} of the diagonal")
        # the penalty's energy, gradient and Hessian pinned before any switch:
        # the first three Newton iterations agree to 1e-5 of the diagonal.  A
        # wrong term moves them by far more; what remains is the linear solve's
        # error in a system whose penalty rows are ~1e6 times stiffer than the
        # rest (at rtol 1e-12: 2.9e-6 with the elastic-and-contact clusters,
        # under 1e-6 with contact-only clusters)
        _, X3o, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=idx,
                                    dbc_target=tgt, max_newton=3)
        d3 = driver.TetAnimFix(sc, pcg_rtol=1e-12)
        d3.ctx.set_dbc_targets(idx, tgt)
        d3.advance_one_time_step(max_newton=3)
        X3, _ = d3.state()
        err3 = np.abs(X3 - X3o).max() / sc.diag
        say(f"[mdbc] after 3 Newton iterations: {err3:.2e} of the diagonal")
        assert err3 < 1e-5
        assert abs(it - it_o) <= 6
        assert err < 2e-3
        move = np.abs(tgt - sc.X[idx]).max()
        assert np.abs(X[idx] - tgt).max() <= 0.01 * move


def test_newton_direction_with_prereduced_partials(monkeypatch):
    """The PCG sums formed once per kernel boundary (the path of meshes with
    at least 1024 update blocks, forced here on the bend bar) reach the same
    direction as the sparse direct solve."""
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla

    monkeypatch.setenv("IDP_PCG_PREREDUCE", "1")
    sc = scenes.config1()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    ctx = capi.Context(0)
    try:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        X = (sc.X + 0.4 * (sc.target - sc.X)).ravel()
        say("[bend] newton direction, pre-reduced PCG partials")
        g, sol, rel, it = ctx.newton_direction(X, sc.target.ravel(), sc.h, dHat2, kappa[0], pcg_rtol=1e-12)
        assert it > 0 and rel <= 1e-12
        rp, col, val = ctx.get_bsr()
        n = rp.size - 1
        A = sps.bsr_matrix((val, col, rp), shape=(3 * n, 3 * n)).tocsc()
        d = spla.spsolve(A, -g)
        assert np.linalg.norm(sol - d) <= 1e-8 * np.linalg.norm(d)
    finally:
        ctx.close()
