"""A compiled C++ program (tests/cpp/step_driver.cpp, no Python in its loop)
drives the step through the C ABI and idp_amd::StepSession on MI355X, in the
Hinge and the Flow variant (DISCRETE_SHELL.h:1110, 1113), on tets and on a
shell with membrane and hinge bending, against the oracle: same Newton
counts, end state within 1e-6 of the bounding-box diagonal.  The program
calls set_mesh before each of its two steps; the mesh must go to the device
once."""
import os
import subprocess

import numpy as np
import pytest

from idp_amd import capi, scenes
from oracle import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "idp_amd", "_lib", "idp_step_driver")


def say(*a):
    print(*a, flush=True)


def _write_input(path, sc, v, b, flow, steps, dHat2, kappa, dbc=None, tgt=None, zero_v=False, shell=None):
    """shell: None, or (X, tris, tri_vol, tri_lam, tri_mu, tri_IB, hinges,
    hinge_rest, bend_k, mass) for a shell-only mesh (sc then gives h and the
    Newton tolerance only)."""
    if shell is None:
        X = sc.X
        vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
        tets = sc.tets
        nE = nH = 0
        bend_k = 0.0
    else:
        X, tris, tvol, tlam, tmu, tIB, hinges, hrest, bend_k, mass = shell
        tets = np.zeros((0, 4), np.int32)
        IB = vol = lam = mu = np.zeros(0)
        nE, nH = tris.shape[0], hinges.shape[0]
    nV, nT = X.shape[0], tets.shape[0]
    dbc = np.zeros(0, np.int32) if dbc is None else np.asarray(dbc, np.int32)
    tgt = np.zeros((0, 3)) if tgt is None else np.asarray(tgt, np.float64)
    with open(path, "wb") as f:
        np.array([nV, nT, dbc.size, int(flow), int(b is not None), steps, int(zero_v), nE, nH], np.int32).tofile(f)
        np.array([sc.h, sc.newton_tol, dHat2, kappa[0], kappa[1], kappa[2], 0.0, bend_k], np.float64).tofile(f)
        for a in (X, v, X):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(tets, np.int32).tofile(f)
        for a in (IB, vol, lam, mu, mass):
            np.ascontiguousarray(a, np.float64).tofile(f)
        if b is not None:
            np.ascontiguousarray(b, np.float64).tofile(f)
        dbc.tofile(f)
        np.ascontiguousarray(tgt, np.float64).tofile(f)
        if shell is not None:
            np.ascontiguousarray(tris, np.int32).tofile(f)
            for a in (tvol, tlam, tmu, tIB):
                np.ascontiguousarray(a, np.float64).tofile(f)
            np.ascontiguousarray(hinges, np.int32).tofile(f)
            np.ascontiguousarray(hrest, np.float64).tofile(f)
    return mass


def _run(tmp_path, name, nV, steps):
    inp, out = tmp_path / f"{name}.in", tmp_path / f"{name}.out"
    say(f"[cpp] {name}: {DRIVER}")
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=240)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    ints = np.frombuffer(raw[: 4 * (steps + 1)], np.int32)
    dbl = np.frombuffer(raw[4 * (steps + 1):], np.float64)
    X = dbl[: 3 * nV].reshape(-1, 3)
    v = dbl[3 * nV: 6 * nV].reshape(-1, 3)
    return ints[:steps], int(ints[steps]), X, v, dbl[6 * nV:]


def test_driver_program_is_built():
    assert os.access(DRIVER, os.X_OK), "build() compiles tests/cpp/step_driver.cpp"


def test_cpp_driver_hinge_two_steps_config1(tmp_path):
    """The 20x4x4 bend bar: a hinge step toward the self-intersecting target,
    then a second step from its state at rest (velocities zeroed between
    steps, as an animation-fix driver does before loading the next frame)."""
    sc = scenes.config1()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    _write_input(tmp_path / "hinge.in", sc, sc.velocity, None, False, 2, dHat2, kappa, zero_v=True)
    iters, uploads, X, v, k3 = _run(tmp_path, "hinge", sc.X.shape[0], 2)
    it1, X1, v1, k1, _ = m.advance(sc.X, sc.velocity, sc.h, sc.newton_tol, dHat2, kappa)
    it2, X2, v2, k2, _ = m.advance(X1, np.zeros_like(v1), sc.h, sc.newton_tol, dHat2, k1)
    err = np.abs(X - X2).max() / sc.diag
    say(f"[cpp] hinge: newton {list(iters)} (oracle {it1}, {it2}), uploads {uploads}, err/diag {err:.2e}")
    assert list(iters) == [it1, it2]
    assert uploads == 1
    assert err < 1e-6
    assert k3[0] == k2[0]


def test_cpp_driver_flow_in_contact(tmp_path):
    """Flow variant with a body force pressing one block into another."""
    sc, _ = scenes.two_blocks_slide()
    m = pyoracle.Mesh(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2, kappa = pyoracle.initialize_oipc(sc.dhat, m.mass)
    nA = sc.X.shape[0] // 2
    acc = np.zeros_like(sc.X)
    acc[nA:, 0] = -0.3
    acc[nA:, 1] = -1.0
    b = m.mass[:, None] * acc
    v0 = np.zeros_like(sc.X)
    _write_input(tmp_path / "flow.in", sc, v0, b, True, 1, dHat2, kappa)
    iters, uploads, X, v, k3 = _run(tmp_path, "flow", sc.X.shape[0], 1)
    it_o, Xo, vo, ko, st = m.advance(sc.X, v0, sc.h, sc.newton_tol, dHat2, kappa, body_force=b, flow=True)
    err = np.abs(X - Xo).max() / sc.diag
    say(f"[cpp] flow: newton {iters[0]} (oracle {it_o}), constraints at the end {st['n_constraints']}, "
        f"err/diag {err:.2e}")
    assert iters[0] == it_o > 1
    assert uploads == 1
    assert err < 1e-6


def test_cpp_driver_shell_hinge_draped(tmp_path):
    """The Hinge variant on a shell: a sheet under gravity draped over a
    sphere shell held by Dirichlet rows (node, x, y, z) at its rest
    positions, membrane and hinge bending through ShellPart, two steps."""
    from types import SimpleNamespace

    rho, young, nu, thick, h = 100.0, 1e4, 0.3, 0.003, 0.01
    X, T, dbc = scenes.draped_sheet()
    tris, mass, tvol, lam, mu = capi.shell_init(X, T, rho, young, nu, thick)
    hinges, hrest, tIB = capi.shell_hinges(X, tris)
    bend_k = young * thick ** 3 / (24 * (1 - nu * nu))
    dHat2, kappa = capi.initialize_oipc(1e-3, mass)
    b = mass[:, None] * np.array([0.0, -9.81, 0.0])[None, :]
    sc = SimpleNamespace(h=h, newton_tol=1e-3 * h)
    v0 = np.zeros_like(X)
    _write_input(tmp_path / "shell.in", sc, v0, b, False, 2, dHat2, kappa, dbc=dbc, tgt=X[dbc],
                 shell=(X, tris, tvol, lam, mu, tIB, hinges, hrest, bend_k, mass))
    iters, uploads, Xg, vg, k3 = _run(tmp_path, "shell", X.shape[0], 2)
    m = pyoracle.ShellMesh(X, T, rho, thick, E=young, nu=nu, bend_mult=1.0)
    m.set_dbc(dbc)
    ko = kappa.copy()
    Xo, vo, its = X.copy(), v0.copy(), []
    for _ in range(2):
        it, Xo, vo, ko, _ = m.advance_hinge(Xo, vo, h, 1e-3 * h, dHat2, ko, dbc_idx=dbc, body_force=b)
        its.append(it)
    diag = float(np.linalg.norm(X.max(0) - X.min(0)))
    err = np.abs(Xg - Xo).max() / diag
    say(f"[cpp] shell: newton {list(iters)} (oracle {its}), uploads {uploads}, err/diag {err:.2e}")
    assert list(iters) == its
    assert uploads == 1
    assert err < 1e-6
    assert np.abs(Xg[dbc] - X[dbc]).max() == 0.0
