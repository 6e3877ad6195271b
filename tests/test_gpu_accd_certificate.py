"""The no-hit certificate of long additive-CCD pairs on MI355X: on the long
pairs logged from the 64x16x16 twist-and-fold bar (hits, no hits and a pair
the step cap stopped) and on a point sliding 1e-8 above a triangle, the device
step size with the certificate, and without it (every pair stepped, no step
cap), equals the oracle's bit for bit (CCD.h:280-395 steps every pair), and
the certificate ends pairs early in these scenes."""
import numpy as np
import pytest

from idp_amd import capi
from oracle import pyoracle
from tests import long_pairs
from tests.conftest import ORACLE_ACCD_CAP

pytestmark = pytest.mark.gpu


@pytest.fixture
def uncapped_oracle():
    pyoracle.set_accd_max_steps(0)  # the reference's loop: no step cap
    yield
    pyoracle.set_accd_max_steps(ORACLE_ACCD_CAP)


def say(*a):
    print(*a, flush=True)


def test_long_pair_step_sizes_match_oracle_with_and_without_certificate(uncapped_oracle):
    cases = long_pairs.load() + [long_pairs.sliding_point(s) for s in (0.5, 0.0, -0.5, -1.9)]
    certified = 0
    for i, pr in enumerate(cases):
        X, T, P = long_pairs.scene(pr)
        m = pyoracle.ShellMesh(X, T, 100.0, 0.003, E=1e4, nu=0.3)
        a_o, nc_o = m.ccd(X, P, pr["alpha"])
        tris, mass, tvol, _, _ = capi.shell_init(X, T, 100.0, 1e4, 0.3, 0.003)
        ctx = capi.Context(0)
        try:
            ctx.set_mesh_shell(X, mass, tris, tvol)
            ctx.set_accd_max_steps(-1)
            say(f"[accd certificate] case {i} ({'EE' if pr['ee'] else 'PT'}, bound {pr['alpha']}): certificate on")
            a_c, nc = ctx.ccd(X, P, pr["alpha"])
            _, cert = ctx.ccd_last_stats()
            say(f"[accd certificate] case {i}: alpha {a_c!r} (oracle {a_o!r}), certified {cert}; certificate off")
            ctx.set_accd_certificate(False)
            a_s, _ = ctx.ccd(X, P, pr["alpha"])
            _, cert_off = ctx.ccd_last_stats()
            say(f"[accd certificate] case {i}: stepped alpha {a_s!r}")
        finally:
            ctx.close()
        assert nc == nc_o
        assert a_c == a_o
        assert a_s == a_o
        assert cert_off == 0
        certified += cert
    assert certified >= 5


@pytest.mark.parametrize("speed", [0.0, -1.9])
def test_sliding_point_with_thickness_matches_oracle(uncapped_oracle, speed):
    """The certificate's gap with a thickness: the loop stops where the
    distance less the thickness falls below a tenth of its start value."""
    th = 1e-4
    pr = long_pairs.sliding_point(speed, thickness=th)
    X, T, P = long_pairs.scene(pr)
    m = pyoracle.ShellMesh(X, T, 100.0, 0.003, E=1e4, nu=0.3)
    a_o, nc_o = m.ccd(X, P, 1.0, th)
    tris, mass, tvol, _, _ = capi.shell_init(X, T, 100.0, 1e4, 0.3, 0.003)
    with capi.Context(0) as ctx:
        ctx.set_mesh_shell(X, mass, tris, tvol)
        ctx.set_accd_max_steps(-1)
        say(f"[accd certificate] sliding point, thickness {th}, normal speed {speed}")
        a_c, nc = ctx.ccd(X, P, 1.0, th)
        _, cert = ctx.ccd_last_stats()
        say(f"[accd certificate] alpha {a_c!r} (oracle {a_o!r}), certified {cert}")
    assert nc == nc_o and a_c == a_o
    if speed == 0.0:
        assert a_c == 1.0 and cert >= 1
    else:
        assert 0.46 < a_c < 0.9 / 1.9
