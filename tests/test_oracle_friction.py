"""CPU pins of the oracle's lagged friction (oracle/friction.cpp, restating
Library/FEM/FRICTION.h and FRICTION_UTILS.h) against what the mathematics
fixes: which constraints carry friction and in what order, the closest
points as orthogonal projections, orthonormal tangent bases orthogonal to the
contact normal, the normal force as the barrier's force on the contact, the
gradient and Hessian as derivatives of the energy (central differences), the
Coulomb magnitude mu N of a sliding contact with zero net force, the
per-component coefficients, and a block on a tilted slab that slides with
acceleration g (sin t - mu cos t) or sticks; the shell step with friction
is the frictionless one as mu -> 0 and slows a sheet sliding over a sphere."""

import numpy as np
import pytest

from idp_amd import scenes
from oracle import pyoracle

DHAT2, KAPPA = 1e-2, 1e3


def _configs(rng, d=0.05):
    """Four contacts (PT, EE, PE with multiplicity 2, PP) between disjoint
    vertex groups, each at distance d, plus a mollified EE entry."""
    X = []
    cs = []
    # PT: p above triangle (t0, t1, t2) in the plane z = 0
    t = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.1, 1.2, 0.0]]) + rng.normal(0, 0.02, (3, 3)) * [1, 1, 0]
    p = np.array([0.3, 0.35, d])
    X += [p, *t]
    cs.append([-0 - 1, 1, 2, 3])
    # EE: edge a along x at z = d over edge b along y
    o = np.array([5.0, 0.0, 0.0])
    a0, a1 = o + [-0.5, 0.1, d], o + [0.6, -0.1, d]
    b0, b1 = o + [0.05, -0.7, 0.0], o + [-0.1, 0.5, 0.0]
    X += [a0, a1, b0, b1]
    cs.append([4, 5, 6, 7])
    # PE: point over an edge, multiplicity 2
    o = np.array([10.0, 0.0, 0.0])
    X += [o + [0.2, 0.0, d], o + [-0.6, 0.0, 0.0], o + [0.7, 0.05, 0.0]]
    cs.append([-8 - 1, 9, 10, -2])
    # PP
    o = np.array([15.0, 0.0, 0.0])
    X += [o + [0.01, 0.02, d], o]
    cs.append([-11 - 1, 12, -1, -1])
    # a mollified EE entry (no friction): reuses the EE vertices
    cs.insert(1, [4, 5, -6 - 1, 7])
    return np.array(X), np.array(cs, np.int32)


def _barrier_gradient(X, cs, thickness=0.0):
    g = np.zeros(X.size)
    cs = np.ascontiguousarray(cs, np.int32).ravel()
    pyoracle.lib().orc_barrier_gradient(np.ascontiguousarray(X.ravel()), np.ascontiguousarray(X.ravel()),
                                        cs.size // 4, cs, DHAT2, KAPPA, thickness, X.shape[0], g)
    return g.reshape(-1, 3)


def test_friction_set_skips_mollified_and_keeps_order():
    X, cs = _configs(np.random.default_rng(0))
    fs = pyoracle.FrictionSet(X, cs, DHAT2, KAPPA)
    assert fs.n == 4
    assert np.array_equal(fs.fcs.reshape(-1, 4), cs[[0, 2, 3, 4]])


@pytest.mark.parametrize("seed", range(5))
def test_closest_points_and_tangent_bases(seed):
    X, cs = _configs(np.random.default_rng(seed))
    X = X + np.random.default_rng(100 + seed).normal(0, 1e-3, X.shape)
    fs = pyoracle.FrictionSet(X, cs, DHAT2, KAPPA)
    cp = fs.cp.reshape(-1, 2)
    B = fs.tangent_basis()
    for k in range(4):
        assert np.abs(B[k].T @ B[k] - np.eye(2)).max() < 1e-14
    # PT: p - closest point is orthogonal to both triangle edges
    p, t0, t1, t2 = X[0], X[1], X[2], X[3]
    r = p - (t0 + cp[0, 0] * (t1 - t0) + cp[0, 1] * (t2 - t0))
    assert abs(r @ (t1 - t0)) < 1e-13 and abs(r @ (t2 - t0)) < 1e-13
    n = np.cross(t1 - t0, t2 - t0)
    assert np.abs(B[0].T @ n).max() < 1e-13 * np.linalg.norm(n)
    # EE: the closest points' difference is orthogonal to both edges
    a0, a1, b0, b1 = X[4], X[5], X[6], X[7]
    r = (a0 + cp[1, 0] * (a1 - a0)) - (b0 + cp[1, 1] * (b1 - b0))
    assert abs(r @ (a1 - a0)) < 1e-13 and abs(r @ (b1 - b0)) < 1e-13
    n = np.cross(a1 - a0, b1 - b0)
    assert np.abs(B[1].T @ n).max() < 1e-13 * np.linalg.norm(n)
    # PE: p - closest point is orthogonal to the edge and to both columns
    p, e0, e1 = X[8], X[9], X[10]
    r = p - (e0 + cp[2, 0] * (e1 - e0))
    assert abs(r @ (e1 - e0)) < 1e-13
    assert np.abs(B[2].T @ r).max() < 1e-13
    # PP: both columns orthogonal to q - p
    assert np.abs(B[3].T @ (X[12] - X[11])).max() < 1e-15


@pytest.mark.parametrize("thickness", [0.0, 0.01])
def test_normal_force_is_the_barrier_force_on_the_contact(thickness):
    X, cs = _configs(np.random.default_rng(1))
    fs = pyoracle.FrictionSet(X, cs, DHAT2, KAPPA, thickness)
    g = _barrier_gradient(X, cs[[0]], thickness)
    assert fs.nforce[0] == pytest.approx(np.linalg.norm(g[0]), rel=1e-12)
    g = _barrier_gradient(X, cs[[2]], thickness)
    assert fs.nforce[1] == pytest.approx(np.linalg.norm(g[4] + g[5]), rel=1e-12)
    g = _barrier_gradient(X, np.array([[-8 - 1, 9, 10, -1]]), thickness)  # multiplicity 1
    assert fs.nforce[2] == pytest.approx(np.linalg.norm(g[8]), rel=1e-12)
    g = _barrier_gradient(X, cs[[4]], thickness)
    assert fs.nforce[3] == pytest.approx(np.linalg.norm(g[11]), rel=1e-12)
    assert np.all(fs.nforce > 0)


def _assemble(H, vid, nV):
    A = np.zeros((3 * nV, 3 * nV))
    for h, v in zip(H, vid):
        nv = int((v >= 0).sum())
        n = 3 * nv
        M = h[: n * n].reshape(n, n)
        idx = np.concatenate([3 * v[a] + np.arange(3) for a in range(nv)])
        A[np.ix_(idx, idx)] += M
    return A


@pytest.mark.parametrize("regime", ["sliding", "static", "mixed"])
def test_gradient_and_hessian_are_derivatives_of_the_energy(regime):
    rng = np.random.default_rng(7)
    X0, cs = _configs(rng)
    fs = pyoracle.FrictionSet(X0, cs, DHAT2, KAPPA)
    epsvh2, mu = 1e-6, 0.4
    scale = {"sliding": 1e-2, "static": 3e-4, "mixed": 1e-3}[regime]
    X = X0 + rng.normal(0, scale, X0.shape)
    E0 = fs.energy(X, X0, epsvh2, mu)
    g = fs.gradient(X, X0, epsvh2, mu)
    fd = np.zeros_like(g)
    step = 1e-7
    for i in range(g.size):
        Xp, Xm = X.ravel().copy(), X.ravel().copy()
        Xp[i] += step
        Xm[i] -= step
        fd[i] = (fs.energy(Xp, X0, epsvh2, mu) - fs.energy(Xm, X0, epsvh2, mu)) / (2 * step)
    assert E0 > 0
    assert np.abs(g - fd).max() <= 1e-6 * np.abs(g).max()
    H, vid = fs.hessian_local(X, X0, epsvh2, mu, project=False)
    Hp, _ = fs.hessian_local(X, X0, epsvh2, mu, project=True)
    assert np.abs(H - Hp).max() <= 1e-12 * np.abs(H).max()  # the inner 2 x 2 is PSD in every regime
    A = _assemble(H, vid, X.shape[0])
    fdH = np.zeros_like(A)
    step = 1e-8
    for i in range(g.size):
        Xp, Xm = X.ravel().copy(), X.ravel().copy()
        Xp[i] += step
        Xm[i] -= step
        fdH[:, i] = (fs.gradient(Xp, X0, epsvh2, mu) - fs.gradient(Xm, X0, epsvh2, mu)) / (2 * step)
    assert np.abs(A - fdH).max() <= 1e-5 * np.abs(A).max()


def test_hessian_at_rest_is_f1_times_tt():
    """No displacement: W = mu N f1(0) I = mu N (2 / epsv h) I, so the PT block
    (p, p) is mu N 2 / epsvh B B^T (the reference's relDXNorm == 0 branch)."""
    X, cs = _configs(np.random.default_rng(2))
    fs = pyoracle.FrictionSet(X, cs, DHAT2, KAPPA)
    epsvh2, mu = 1e-6, 0.3
    H, _ = fs.hessian_local(X, X, epsvh2, mu)
    B = fs.tangent_basis()[0]
    want = mu * fs.nforce[0] * 2 / np.sqrt(epsvh2) * B @ B.T
    assert np.abs(H[0].reshape(12, 12)[:3, :3] - want).max() <= 1e-12 * np.abs(want).max()


def test_sliding_contact_has_coulomb_force_and_no_net_force():
    X0, cs = _configs(np.random.default_rng(3))
    fs = pyoracle.FrictionSet(X0, cs, DHAT2, KAPPA)
    epsvh2, mu = 1e-8, 0.25
    B = fs.tangent_basis()
    X = X0.copy()
    X[0] += B[0] @ np.array([3e-3, -2e-3])  # PT point slides in the tangent plane
    X[8] += B[2] @ np.array([1e-3, 4e-3])   # PE point
    g = fs.gradient(X, X0, epsvh2, mu).reshape(-1, 3)
    assert np.linalg.norm(g[0]) == pytest.approx(mu * fs.nforce[0], rel=1e-12)
    u = B[0].T @ (X[0] - X0[0])
    assert np.allclose(g[0] / np.linalg.norm(g[0]), B[0] @ u / np.linalg.norm(u), rtol=0, atol=1e-12)
    assert np.linalg.norm(g[8]) == pytest.approx(2 * mu * fs.nforce[2], rel=1e-12)  # multiplicity 2
    for grp in ([0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10], [11, 12]):
        assert np.abs(g[grp].sum(0)).max() <= 1e-15 * max(np.abs(g).max(), 1.0)


def test_point_point_friction_reads_the_second_points_position():
    """FRICTION.h:212-214: the PP relative displacement is dx_p - x_q."""
    X0, cs = _configs(np.random.default_rng(4))
    fs = pyoracle.FrictionSet(X0, cs[[4]], DHAT2, KAPPA)
    epsvh, mu = 1e-3, 0.5
    X = X0.copy()
    X[11] += [1e-3, 2e-3, 0.0]
    B = fs.tangent_basis()[0]
    u = B.T @ ((X[11] - X0[11]) - X[12])
    x2 = u @ u
    f0 = np.sqrt(x2) if x2 >= epsvh ** 2 else x2 * (-np.sqrt(x2) / 3 + epsvh) / epsvh ** 2 + epsvh / 3
    assert fs.energy(X, X0, epsvh ** 2, mu) == pytest.approx(mu * fs.nforce[0] * f0, rel=1e-14)
    # moving q and its start together (no displacement of q) changes the energy
    X2, X02 = X.copy(), X0.copy()
    X2[12] += [0.5, 0.0, 0.0]
    X02[12] += [0.5, 0.0, 0.0]
    assert fs.energy(X2, X02, epsvh ** 2, mu) != fs.energy(X, X0, epsvh ** 2, mu)


def test_coef_scales_normal_forces_by_component_pair():
    X, cs = _configs(np.random.default_rng(5))
    fs = pyoracle.FrictionSet(X, cs, DHAT2, KAPPA)
    N0 = fs.nforce.copy()
    comp_range = np.array([6, 11, 13], np.int32)  # components [0, 6), [6, 11), [11, 13)
    mu_comp = np.arange(1, 10, dtype=np.float64) / 10
    fs.coef(comp_range, mu_comp)
    # PT (p 0, t0 1): (0, 0); EE (a0 4, b0 6): (0, 1); PE (p 8, e0 9): (1, 1); PP (p 11, q 12): (2, 2)
    want = N0 * mu_comp[[0 + 0 * 3, 0 + 1 * 3, 1 + 1 * 3, 2 + 2 * 3]]
    assert np.array_equal(fs.nforce, want)
    with pytest.raises(ValueError):
        pyoracle.FrictionSet(X, cs, DHAT2, KAPPA).coef(np.array([6, 11], np.int32), mu_comp[:4])


def _slab_run(mu, steps, fric_iters, theta=20.0):
    sc, ns, dbc = scenes.block_on_slab()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    th = np.radians(theta)
    b = m.mass[:, None] * (9.81 * np.array([np.sin(th), -np.cos(th), 0.0]))[None, :]
    X, v, k = sc.X.copy(), np.zeros_like(sc.X), kappa.copy()
    vx, sts = [], []
    for _ in range(steps):
        it, X, v, k, st = m.advance_friction(X, v, sc.h, sc.newton_tol, dHat2, k, mu, 1e-6, fric_iters=fric_iters,
                                             dbc_idx=dbc, body_force=b)
        vx.append(v[ns:, 0].mean())
        sts.append(st)
    return sc, np.array(vx), sts, th


def test_block_slides_with_coulomb_acceleration():
    sc, vx, sts, th = _slab_run(0.1, 8, 0)
    a = 9.81 * (np.sin(th) - 0.1 * np.cos(th))
    a_fit = (vx[7] - vx[3]) / (4 * sc.h)
    assert a_fit == pytest.approx(a, rel=1e-2)
    assert all(st["n_friction"] > 0 for st in sts)
    assert any(st["friction_updates"] > 0 for st in sts)


@pytest.mark.parametrize("fric_iters", [1, 3])
def test_block_sticks_and_updates_stay_within_fric_iters(fric_iters):
    sc, vx, sts, th = _slab_run(0.5, 8, fric_iters)
    free = 9.81 * np.sin(th) * 8 * sc.h  # the frictionless speed after these steps
    # the block rings elastically about rest (a few mm/s) instead of sliding
    assert abs(vx[-1]) < 0.05 * free and np.abs(vx).max() < 0.1 * free
    assert all(st["friction_updates"] <= fric_iters - 1 for st in sts)


def _sheet_run(mu, steps=8, fric_iters=1):
    RHO, YOUNG, NU, THICK, H = 100.0, 1e4, 0.3, 0.003, 0.01
    X, T, dbc = scenes.draped_sheet()
    m = pyoracle.ShellMesh(X, T, RHO, THICK, E=YOUNG, nu=NU, bend_mult=1.0)
    m.set_dbc(dbc)
    bf = m.mass[:, None] * np.array([0.0, -9.81, 0.0])[None, :]
    dHat2, kappa = pyoracle.initialize_oipc(1e-3, m.mass)
    Xs, v = X.copy(), np.zeros_like(X)
    sts = []
    for _ in range(steps):
        if mu is None:
            _, Xs, v, kappa, st = m.advance_hinge(Xs, v, H, 1e-3 * H, dHat2, kappa, dbc_idx=dbc, body_force=bf)
        else:
            _, Xs, v, kappa, st = m.advance_hinge_friction(Xs, v, H, 1e-3 * H, dHat2, kappa, mu, 1e-6,
                                                           fric_iters=fric_iters, dbc_idx=dbc, body_force=bf)
        sts.append(st)
    return X, dbc, Xs, v, sts


def test_shell_friction_vanishes_with_mu_and_slows_the_draped_sheet():
    """The shell step with friction (Advance_One_Step_IE_Hinge with lagged
    friction): with mu -> 0 it is the frictionless step, and with mu = 2 the
    sheet's vertices on the sphere slide over it more slowly."""
    X, dbc, X0, v0, _ = _sheet_run(None, steps=10)
    _, _, Xe, ve, sts = _sheet_run(1e-12, steps=10)
    assert sum(st["n_friction"] for st in sts) > 0
    diag = float(np.linalg.norm(X.max(0) - X.min(0)))
    assert np.abs(Xe - X0).max() <= 1e-9 * diag
    _, _, X2, v2, sts2 = _sheet_run(2.0, steps=10)
    assert sts2[-1]["n_friction"] > 0
    c = np.array([0.5, -0.303, 0.5])  # the sphere's centre (radius 0.3)
    free = np.setdiff1d(np.arange(X.shape[0]), dbc)

    def tangential(Xs, v):
        r = Xs[free] - c
        d = np.linalg.norm(r, axis=1)
        n = r / d[:, None]
        vt = v[free] - (v[free] * n).sum(1)[:, None] * n
        return np.linalg.norm(vt[d < 0.31], axis=1)

    t0, t2 = tangential(X0, v0), tangential(X2, v2)
    assert t0.size > 0 and t2.size == t0.size
    assert t2.max() < 0.95 * t0.max() and t2.mean() < 0.95 * t0.mean()
