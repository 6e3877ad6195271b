"""Stage-by-stage parity on the 160x48x48 crumple bar (386,561 vertices,
2,211,840 tets, 70,656 boundary triangles) on MI355X against the oracle:

- the ACCD step size along the crumple direction from the jittered rest
  state (4.3M swept-box candidates), bit for bit;
- the constraint set at 90% of that step with a contact radius of 10 dhat
  (about 1M constraints), bit for bit after canonical sort;
- barrier energy (1e-10 relative) and gradient (1e-10 in norm) on that set;
- the projected barrier Hessian of every 5th constraint of that set (every
  kind, ~200k constraints) within 1e-9 of each row's largest entry, with the
  order-9 kinds on lane triples and on 16-lane teams;
- the Newton direction there: the gradient against the oracle's elastic +
  inertia + barrier gradient (1e-10), and the PCG solution's residual
  recomputed on the host from the device's own BSR matrix.
"""
import numpy as np
import pytest

from idp_amd import capi, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _canon(cs):
    cs = np.ascontiguousarray(cs, dtype=np.int64).reshape(-1, 4)
    order = np.lexsort((cs[:, 3], cs[:, 2], cs[:, 1], cs[:, 0]))
    return cs[order]


class Crumple:
    def __init__(self):
        self.sc = sc = scenes.config5()
        say(f"[crumple] {sc.X.shape[0]} vertices, {sc.tets.shape[0]} tets")
        self.m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        self.dHat2, self.kappa = pyoracle.initialize_oipc(sc.dhat, self.m.mass)
        vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        self.ctx = capi.Context(0)
        self.ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        self.p = (sc.target - sc.X).ravel()
        self.alpha_o, self.nc_o = self.m.ccd(sc.X, self.p, 1.0)
        self.X1 = sc.X.ravel() + 0.9 * self.alpha_o * self.p
        self.r2 = 100.0 * self.dHat2


@pytest.fixture(scope="module")
def crumple():
    c = Crumple()
    yield c
    c.ctx.close()


def test_crumple_surface_sizes(crumple):
    assert crumple.ctx.surface_sizes() == (35330, 105984, 70656)


def test_crumple_ccd_step_bit_exact(crumple):
    c = crumple
    a, nc = c.ctx.ccd(c.sc.X, c.p, 1.0)
    say(f"[crumple] ccd alpha {a!r} (oracle {c.alpha_o!r}), candidates {nc} (oracle {c.nc_o})")
    assert 0 < a < 1
    assert a == c.alpha_o


def test_crumple_constraint_set_bit_exact(crumple):
    c = crumple
    g = c.ctx.constraint_set(c.X1, c.r2)
    o = c.m.constraint_set(c.X1, c.r2)
    say(f"[crumple] constraint set: {g.shape[0]} (oracle {o.shape[0]})")
    assert o.shape[0] > 500000
    assert np.array_equal(_canon(g), _canon(o))


def test_crumple_barrier_energy_and_gradient(crumple):
    c = crumple
    cs = c.m.constraint_set(c.X1, c.r2)
    k = c.kappa[0]
    Eg = c.ctx.barrier_energy(c.X1, cs, c.r2, k)
    Eo = c.m.barrier_energy(c.X1, cs, c.r2, k)
    gg = c.ctx.barrier_gradient(c.X1, cs, c.r2, k)
    go = c.m.barrier_gradient(c.X1, cs, c.r2, k)
    say(f"[crumple] barrier E {Eg!r} (oracle {Eo!r}), |dg|/|g| {np.linalg.norm(gg - go) / np.linalg.norm(go):.2e}")
    assert abs(Eg - Eo) <= 1e-10 * abs(Eo)
    assert np.linalg.norm(gg - go) <= 1e-10 * np.linalg.norm(go)


def test_crumple_barrier_hessian_sample(crumple):
    c = crumple
    cs = c.m.constraint_set(c.X1, c.r2)[::5].copy()
    k = c.kappa[0]
    Ho, vo = c.m.barrier_hessian_local(c.X1, cs, c.r2, k, project=True)
    scale = np.abs(Ho).max(1)
    team0 = c.ctx.barrier_team()
    try:
        for team in (3, 16):
            c.ctx.set_barrier_team(team)
            say(f"[crumple] projected barrier Hessian of {cs.shape[0]} constraints, team {team}")
            Hg, vg = c.ctx.barrier_hessian_local(c.X1, cs, c.r2, k, project=True)
            assert np.array_equal(vg, vo)
            err = (np.abs(Hg - Ho).max(1) / np.where(scale > 0, scale, 1.0)).max()
            say(f"[crumple] team {team}: max row error / row max {err:.2e}")
            assert err <= 1e-9
            del Hg
    finally:
        c.ctx.set_barrier_team(team0)


def test_crumple_newton_direction(crumple):
    import scipy.sparse as sps

    c = crumple
    sc = c.sc
    xt = sc.target.ravel()
    g, sol, rel, it = c.ctx.newton_direction(c.X1, xt, sc.h, c.r2, c.kappa[0], pcg_rtol=1e-8, pcg_maxit=40000)
    say(f"[crumple] newton direction: {it} PCG iterations, relative residual {rel:.2e}")
    assert 0 < it < 40000 and rel <= 1e-8
    m = c.m
    cs = m.constraint_set(c.X1, c.r2)
    go = m.elastic_gradient(c.X1, sc.h ** 2) + np.repeat(m.mass, 3) * (c.X1 - xt)
    go += m.barrier_gradient(c.X1, cs, c.r2, c.kappa[0])
    assert np.linalg.norm(g - go) <= 1e-10 * np.linalg.norm(go)
    rp, col, val = c.ctx.get_bsr()
    n = rp.size - 1
    A = sps.bsr_matrix((val, col, rp), shape=(3 * n, 3 * n)).tocsr()
    true_rel = np.linalg.norm(A @ sol + g) / np.linalg.norm(g)
    say(f"[crumple] host-recomputed residual {true_rel:.2e}, {col.size} blocks")
    assert true_rel <= 2e-8
