"""Pins the oracle against the sympy golden vectors (CPU).

The golden values are symbolic derivatives of the closed-form distances,
barrier, mollifier and Neo-Hookean energy (tests/golden/make_golden.py); the
oracle differentiates the same closed forms with forward-mode jets and
projects with its own Jacobi eigensolver, so a dropped term, a wrong sign or
a transposed operand in either shows up here.
"""
import json
import os

import numpy as np
import pytest

from oracle import pyoracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden_derivatives.json")))


class _Bare:
    """Constraint evaluation on a 4-vertex scratch mesh."""

    def __init__(self, X, X0):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.X0 = np.ascontiguousarray(X0, dtype=np.float64)


def _eval(case, project):
    L = pyoracle.lib()
    import ctypes as C

    X = np.ascontiguousarray(case["X"], dtype=np.float64)
    X0 = np.ascontiguousarray(case["X0"], dtype=np.float64)
    cs = np.ascontiguousarray(case["cs"], dtype=np.int32)
    E = C.c_double()
    assert L.orc_barrier_energy(X, X0, 1, cs, case["dHat2"], case["kappa"], 0.0, C.byref(E), None) == 0
    g = np.zeros(12)
    assert L.orc_barrier_gradient(X, X0, 1, cs, case["dHat2"], case["kappa"], 0.0, 4, g) == 0
    H = np.zeros(144)
    vid = np.zeros(4, np.int32)
    assert L.orc_barrier_hessian_local(X, X0, 1, cs, case["dHat2"], case["kappa"], 0.0, int(project), H, vid) == 0
    n = 3 * case["nv"]
    return E.value, g[:n], H[: n * n].reshape(n, n), vid


@pytest.mark.parametrize("case", GOLD["barrier"], ids=lambda c: c["kind"])
def test_barrier_energy_gradient_hessian(case):
    E, g, H, vid = _eval(case, project=False)
    n = 3 * case["nv"]
    assert list(vid[: case["nv"]]) == list(range(case["nv"]))
    assert abs(E - case["E"]) <= 1e-11 * abs(case["E"])
    ge = np.array(case["g"])
    assert np.linalg.norm(g - ge) <= 1e-10 * np.linalg.norm(ge)
    He = np.array(case["H"])
    assert He.shape == (n, n)
    assert np.abs(H - He).max() <= 1e-10 * np.abs(He).max()


@pytest.mark.parametrize("case", GOLD["barrier"], ids=lambda c: c["kind"])
def test_barrier_hessian_projection(case):
    _, _, H, _ = _eval(case, project=True)
    He = np.array(case["Hproj"])
    assert np.abs(H - He).max() <= 1e-9 * np.abs(np.array(case["H"])).max()
    assert np.linalg.eigvalsh(0.5 * (H + H.T)).min() >= -1e-9 * np.abs(He).max()


@pytest.mark.parametrize("case", GOLD["barrier_scalar"])
def test_barrier_scalar(case):
    # b, b', b'' through a point-point constraint in 1D: d = |p - q|^2
    import ctypes as C

    L = pyoracle.lib()
    d = case["d"]
    X = np.zeros(12)
    X[0] = np.sqrt(d)
    X0 = np.zeros(12)
    cs = np.array([-1, 1, -1, -1], np.int32)
    E = C.c_double()
    assert L.orc_barrier_energy(X, X0, 1, cs, case["dh"], case["k"], 0.0, C.byref(E), None) == 0
    assert abs(E.value - case["b"]) <= 1e-9 * abs(case["b"]) + 1e-300
    # gradient along x: db/dx = b'(d) * 2 sqrt(d)
    g = np.zeros(12)
    assert L.orc_barrier_gradient(X, X0, 1, cs, case["dh"], case["k"], 0.0, 4, g) == 0
    assert abs(g[0] - case["db"] * 2 * X[0]) <= 1e-9 * abs(case["db"] * 2 * X[0])


def _single_tet(F):
    """Rest tet (0, e1, e2, e3): IB = I, vol = 1/6; deformed so that Ds = F."""
    Xr = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    F = np.array(F).reshape(3, 3)
    Xd = np.vstack([np.zeros(3), F[:, 0], F[:, 1], F[:, 2]])
    return Xr, Xd


def _dFdx():
    """D (9 x 12): vec index j*3+i of F(i,j) versus node-major DOFs."""
    D = np.zeros((9, 12))
    c = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)  # c[node][j]
    for i in range(3):
        for j in range(3):
            for a in range(4):
                D[j * 3 + i, 3 * a + i] = c[a, j]
    return D


@pytest.mark.parametrize("case", GOLD["neohookean"])
def test_neohookean(case):
    Xr, Xd = _single_tet(case["F"])
    tets = np.array([[0, 1, 2, 3]], np.int32)
    m = pyoracle.Mesh(Xr, tets, 1000.0, 100.0, 0.3)
    m.lam[:] = case["lam"]
    m.mu[:] = case["mu"]
    w = 0.04 ** 2
    vol = 1.0 / 6.0
    assert abs(m.vol[0] - vol) < 1e-15
    assert np.allclose(m.IB.reshape(3, 3), np.eye(3))
    E = m.elastic_energy(Xd, w)
    assert abs(E - w * vol * case["psi"]) <= 1e-10 * abs(w * vol * case["psi"]) + 1e-14 * w * vol * case["mu"]
    # gradient: golden P is d psi / dF(i,j), row-major F index 3i+j
    P = np.array(case["P"]).reshape(3, 3)
    g = m.elastic_gradient(Xd, w).reshape(4, 3)
    ge = np.zeros((4, 3))
    for k in range(3):
        ge[k + 1] = w * vol * P[:, k]
    ge[0] = -ge[1:].sum(0)
    assert np.abs(g - ge).max() <= 1e-10 * np.abs(ge).max()
    # Hessian: golden H over row-major F entries -> reorder to vec index j*3+i
    perm = [3 * i + j for j in range(3) for i in range(3)]
    H9 = np.array(case["H"])[np.ix_(perm, perm)]
    H9p = np.array(case["Hproj"])[np.ix_(perm, perm)]
    D = _dFdx()
    for project, ref in ((False, H9), (True, H9p)):
        K = m.elastic_hessian_local(Xd, w, project=project)[0].reshape(12, 12)
        Ke = w * vol * D.T @ ref @ D
        assert np.abs(K - Ke).max() <= 1e-9 * np.abs(Ke).max()


def test_oracle_moving_dbc_paths():
    """Moving DBC in the oracle: a move that fits sets the nodes to their
    targets; one that the CCD cuts short goes through the penalty, which ends
    off with the nodes within 1% of the move; targets equal to the current
    positions give the static-DBC step (the oracle's threaded sums vary in the
    last bits from run to run, so to 1e-14 of the diagonal)."""
    from idp_amd import scenes

    for push, gap in ((0.1, 0.5), (0.5, 0.2)):
        sc, idx, tgt = scenes.two_blocks_push(push=push, gap=gap)
        m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
        it, X, _, _, st = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=idx,
                                    dbc_target=tgt)
        move = np.abs(tgt - sc.X[idx]).max()
        if push < gap:
            assert st["dbc_alpha"] == 1.0
            assert np.array_equal(X[idx], tgt)
        else:
            assert 0.0 < st["dbc_alpha"] < 1.0
            assert st["dbc_stiff"] == 0.0
            assert np.abs(X[idx] - tgt).max() <= 0.01 * move
        # targets at the start positions: the static path
        it_s, Xs, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=idx)
        it_t, Xt, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa, dbc_idx=idx,
                                      dbc_target=sc.X[idx])
        assert it_s == it_t and np.abs(Xs - Xt).max() <= 1e-14 * sc.diag
