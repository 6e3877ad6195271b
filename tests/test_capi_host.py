"""CPU checks of the C ABI library (no device compute).

- libidp_amd.so loads and exports every entry point declared in
  include/idp_amd.h;
- the host-side set-up helpers (volumes/inverse bases/mass, surface
  extraction, kappa initialisation) agree exactly with the oracle.
"""
import os
import re

import numpy as np
import pytest

from idp_amd import capi, scenes
from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "idp_amd.h")).read()
    return sorted(set(re.findall(r"\b(idp_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert sorted(capi.EXPORTED) == names


def test_no_device_is_reported_cleanly():
    # on a host without a GPU the context constructor fails with a status, not a crash
    if capi.device_count() > 0:
        pytest.skip("device present")
    with pytest.raises(capi.IdpError):
        capi.Context(0)


@pytest.mark.parametrize("sc", [scenes.small_bend(), scenes.config1(), scenes.cube_sinusoid(5, seed=3)],
                         ids=lambda s: s.name)
def test_host_setup_matches_oracle(sc):
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    assert np.array_equal(vol, m.vol) and np.array_equal(IB, m.IB)
    assert np.array_equal(lam, m.lam) and np.array_equal(mu, m.mu) and np.array_equal(mass, m.mass)
    bn, be, bt = capi.surface(sc.X, sc.tets)
    assert np.array_equal(bn, m.bnode) and np.array_equal(be, m.bedge) and np.array_equal(bt, m.btri)
    d2, k = capi.initialize_oipc(sc.dhat, mass)
    d2o, ko = pyoracle.initialize_oipc(sc.dhat, m.mass)
    assert d2 == d2o
    assert np.allclose(k, ko, rtol=1e-14, atol=0)


def test_kappa_formula_closed_form():
    # Initialize_EIPC<double,false>: kappa0 = 1e11 * mean(m) * 3 / (4e-16 * b''(1e-16; dhat^2)) with kappa=1
    mass = np.full(10, 2.0)
    d2, k = capi.initialize_oipc(1e-3, mass)
    t = 1e-16 - 1e-6
    b2 = -2 * np.log(1e-16 / 1e-6) - 4 * t / 1e-16 + t * t / 1e-32
    assert d2 == 1e-6
    assert abs(k[0] - 1e11 * 2.0 * 3 / (4e-16 * b2)) <= 1e-14 * k[0]
    assert k[1] == 100 * k[0] and k[2] == 0
