"""The coarse level of the one-mesh PCG preconditioner on MI355X (rigid-body
modes of voxel aggregates added to the cluster block Jacobi): the Newton
direction it gives is the one the solve without it gives, to the solves'
relative residual, in far fewer iterations, and Dirichlet rows of the
direction stay exactly zero (the aggregates leave those nodes out)."""
import numpy as np
import pytest

from idp_amd import driver, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_coarse_level_gives_the_same_direction_in_fewer_iterations():
    sc = scenes.config2()
    d = driver.TetAnimFix(sc)
    try:
        Xt = sc.X + sc.h * sc.velocity  # x~ of the step
        # the Newton system after 200 iterations of the twist-and-fold step
        # (folded, in contact: the systems whose solves take thousands of
        # PCG iterations)
        say("[coarse] 64x16x16 bar: 200 Newton iterations")
        d.ctx.advance(sc.h, sc.newton_tol, d.dHat2, d.kappa, max_newton=200, coarse_verts=-1)
        X, _ = d.ctx.get_state()
        say("[coarse] its Newton system without and with the coarse level")
        g0, s0, rel0, it0 = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-10,
                                                   coarse_verts=-1)
        g1, s1, rel1, it1 = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-10,
                                                   coarse_verts=144)
        say(f"[coarse] PCG iterations {it0} -> {it1}")
        assert np.array_equal(g0, g1)
        assert rel0 <= 1e-10 and rel1 <= 1e-10
        assert np.linalg.norm(s1 - s0) <= 1e-8 * np.linalg.norm(s0)
        assert it1 < 0.6 * it0
    finally:
        d.close()


def test_coarse_level_keeps_dirichlet_rows_zero():
    sc = scenes.config1()  # 525 vertices: above the coarse level's size floor
    d = driver.TetAnimFix(sc)
    try:
        dbc = np.where(sc.X[:, 0] < 1e-9)[0].astype(np.int32)
        assert dbc.size > 0
        d.ctx.set_dbc(dbc)
        X = sc.X
        Xt = X + sc.h * sc.velocity
        say(f"[coarse] 20x4x4 bar with {dbc.size} Dirichlet nodes")
        g, s, rel, it = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-10, coarse_verts=48)
        g0, s0, _, it0 = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-10, coarse_verts=-1)
        s = s.reshape(-1, 3)
        assert np.all(s[dbc] == 0.0)
        assert np.linalg.norm(s.ravel() - s0) <= 1e-8 * np.linalg.norm(s0)
        say(f"[coarse] PCG iterations {it0} -> {it}")
    finally:
        d.close()


def test_coarse_level_is_deterministic_and_bounded_on_a_large_mesh():
    """The 160x48x48 crumple bar (384k vertices): the aggregates are capped in
    number (the dense A_c^-1 is read by every PCG iteration), and two solves
    of the same system give the same direction bit for bit."""
    sc = scenes.config5()
    d = driver.TetAnimFix(sc)
    try:
        X = sc.X
        Xt = X + 0.5 * sc.h * sc.velocity
        say("[coarse] 160x48x48 bar: two solves of its first Newton system")
        g0, s0, rel0, it0 = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-8, coarse_verts=144)
        g1, s1, rel1, it1 = d.ctx.newton_direction(X, Xt, sc.h, d.dHat2, d.kappa[0], pcg_rtol=1e-8, coarse_verts=144)
        say(f"[coarse] PCG iterations {it0}, {it1}")
        assert np.array_equal(s0, s1) and it0 == it1
        assert rel0 <= 1e-8
    finally:
        d.close()
