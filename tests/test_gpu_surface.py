"""Boundary extraction on the device (idp_set_mesh) against the oracle's
restatement of Find_Surface_TriMesh + Find_Surface_Primitives_And_Compute_Area
(MESHIO.h:486-545, 769-835): nodes, edges and triangles equal element for
element, in order, on bars, cubes, a batch of uneven cubes, a mesh with
shuffled tet order and vertex order, and the 160x48x48 bar."""
import numpy as np
import pytest

from idp_amd import capi, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _check(X, T, name):
    vol, IB, lam, mu, mass = capi.fem_init(X, T, 1000.0, 1e5, 0.4)
    ctx = capi.Context(0)
    try:
        ctx.set_mesh(X, T, IB, vol, lam, mu, mass)
        bn, be, bt = ctx.get_surface()
    finally:
        ctx.close()
    on, oe, ot = pyoracle.surface(X, T)
    say(f"[surface] {name}: {bn.size} nodes, {be.shape[0]} edges, {bt.shape[0]} triangles")
    assert np.array_equal(bn, on)
    assert np.array_equal(be, oe)
    assert np.array_equal(bt, ot)


@pytest.mark.parametrize("dims", [(1, 1, 1), (8, 2, 2), (20, 4, 4), (5, 7, 3)])
def test_device_surface_bars(dims):
    X, T = scenes.kuhn_block(*dims, 0.05)
    _check(X, T.astype(np.int32), f"block{dims}")


def test_device_surface_shuffled_order():
    """Tets in random order and vertex ids permuted: the map orders depend on
    the ids, not on the input order."""
    X, T = scenes.kuhn_block(6, 5, 4, 0.05)
    rng = np.random.default_rng(5)
    perm = rng.permutation(X.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    X2 = X[perm]
    T2 = inv[T][rng.permutation(T.shape[0])].astype(np.int32)
    _check(X2, T2, "shuffled block")


def test_device_surface_disjoint_cubes():
    parts, off = [], 0
    Xs, Ts = [], []
    for n in (4, 7, 5):
        X, T = scenes.kuhn_block(n, n, n, 0.05)
        Xs.append(X + np.array([off * 0.05, 0.0, 0.0]))
        Ts.append(T + sum(x.shape[0] for x in Xs[:-1]))
        off += n + 2
    _check(np.vstack(Xs), np.vstack(Ts).astype(np.int32), "three cubes")


def test_device_surface_twist_fold_bar():
    sc = scenes.config2()
    _check(sc.X, sc.tets, sc.name)


def test_device_surface_large_bar():
    X, T = scenes.kuhn_block(160, 48, 48, 0.05)
    _check(X, T.astype(np.int32), "160x48x48 bar")


def test_load_velocity_on_device():
    """Load_Velocity_X0 (MESHIO.h:987-1005) on the device: v = (target - x) / h,
    bit for bit the host expression, from the context's current positions."""
    sc = scenes.small_bend()
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    ctx = capi.Context(0)
    try:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        X = sc.X + 0.01 * np.sin(sc.X)
        ctx.set_state(X, np.zeros_like(X))
        for k in range(2):  # the second call reuses the staging buffer
            tgt = sc.target + 0.1 * k
            ctx.load_velocity(tgt, sc.h)
            Xg, v = ctx.get_state()
            assert np.array_equal(Xg, X)
            assert np.array_equal(v, (tgt - X) / sc.h)
    finally:
        ctx.close()
