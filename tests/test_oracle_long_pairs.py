"""CPU pins of the long additive-CCD scenes (tests/long_pairs.py) on the
oracle's step size (Compute_Intersection_Free_StepSize, CCD.h:280-395): a
point sliding 1e-8 above a triangle ends with no hit unless its normal motion
brings it to a tenth of its distance within the step, at t = 0.9 / speed
(less the last conservative step); the fixture pairs of the bar are
candidates of their two-triangle scenes, and those logged as hits still hit."""
import numpy as np
import pytest

from oracle import pyoracle
from tests import long_pairs
from tests.conftest import ORACLE_ACCD_CAP


@pytest.fixture(autouse=True)
def uncapped_oracle():
    pyoracle.set_accd_max_steps(0)  # the reference's loop: no step cap
    yield
    pyoracle.set_accd_max_steps(ORACLE_ACCD_CAP)


def _ccd(pr):
    X, T, P = long_pairs.scene(pr)
    m = pyoracle.ShellMesh(X, T, 100.0, 0.003, E=1e4, nu=0.3)
    return m.ccd(X, P, pr["alpha"])


@pytest.mark.parametrize("speed", [0.5, 0.0, -0.5])
def test_sliding_point_has_no_hit(speed):
    a, nc = _ccd(long_pairs.sliding_point(speed))
    assert nc >= 1 and a == 1.0


def test_approaching_sliding_point_hits_at_a_tenth_of_its_distance():
    a, nc = _ccd(long_pairs.sliding_point(-1.9))
    t = 0.9 / 1.9
    # the last step before the stopping evaluation is at most (1 - eta) d / maxDisp long
    assert t - 0.9 * 0.1e-8 / 1e-3 * 2 < a < t


def test_fixture_pairs_are_candidates_and_logged_hits_hit():
    pairs = long_pairs.load()
    assert {p["result"] for p in pairs} == {1, 2, 3}
    for pr in pairs:
        a, nc = _ccd(pr)
        assert nc >= 1
        assert 0.0 < a <= pr["alpha"]
        if pr["result"] == 1:
            assert a < pr["alpha"]
        if pr["result"] == 3:  # stopped by the device's step cap: no hit when stepped to the end
            assert a == pr["alpha"]
