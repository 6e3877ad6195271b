"""PCG launch configuration on MI355X, through the C ABI.

The update width and the placement of the scalar wavefront do not change
any arithmetic, so results are compared bit for bit across them.  Another
lanes-per-row count G changes the order of each block row's sum, so its
result is compared with the oracle at the usual 1e-6 of the bounding-box
diagonal instead.
"""
import json

import numpy as np
import pytest

from idp_amd import autotune, driver, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _solve(sc, shape, waves):
    d = driver.TetAnimFix(sc)
    d.ctx.set_launch_config(shape, waves)
    assert d.ctx.launch_config() == (shape, waves)
    it, st = d.advance_one_time_step()
    X, v = d.state()
    return it, st["pcg_iters"], X, v


def test_launch_config_same_lanes_per_row_is_bit_identical():
    sc = scenes.cube_sinusoid(5, seed=7)
    say("[tuning] G = 8: scalar wavefront 0 / fifth, update width 1 / 4 / 8")
    ref = _solve(sc, 5, 4)
    for shape, waves in ((1, 4), (5, 1), (5, 8), (1, 2)):
        got = _solve(sc, shape, waves)
        say(f"[tuning] shape {shape} width {waves}: newton {got[0]} pcg {got[1]}")
        assert got[0] == ref[0] and got[1] == ref[1]
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3])


def test_launch_config_other_lanes_per_row_matches_oracle():
    sc = scenes.small_bend()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    it_o, Xo, _, _, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa)
    for shape in (0, 2, 3, 6):
        it, _, X, _ = _solve(sc, shape, 2)
        err = float(np.abs(X - Xo).max()) / sc.diag
        say(f"[tuning] shape {shape}: newton {it} (oracle {it_o}), err/diag {err:.2e}")
        assert err < 1e-6


def test_autotune_writes_a_usable_choice(tmp_path):
    out = tmp_path / "tuning.json"
    say("[tuning] autotune on the small bend, 3 Newton iterations per configuration")
    autotune.main(["--out", str(out), "--workloads", "small_bend", "--newton", "3"])
    d = json.loads(out.read_text())
    assert 0 <= d["spmv_shape"] <= 7 and d["upd_waves"] in autotune.UPD_WAVES
    assert len(d["measurements"]) == len(autotune.SHAPE_TABLE) + len(autotune.UPD_WAVES) - 1
    assert all(r["spmv_us"] > 0 for r in d["measurements"][: len(autotune.SHAPE_TABLE)])
    it, _, _, _ = _solve(scenes.small_bend(), d["spmv_shape"], d["upd_waves"])
    assert it > 0
