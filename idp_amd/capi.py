"""ctypes binding of ``include/idp_amd.h`` (the C ABI of libidp_amd.so).

The HIP library is mandatory: importing this module raises if the shared
library has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libidp_amd.so")

IDP_OK = 0
IDP_E_INVALID = -1
IDP_E_HIP = -2
IDP_E_NONPOS_DIST = -3
IDP_E_NOT_SPD = -4
IDP_E_CCD_ZERO = -5
IDP_E_CAPACITY = -6
IDP_E_INVALID_START = -7
IDP_E_NO_GPU = -8

_STATUS = {
    IDP_E_INVALID: "invalid argument",
    IDP_E_HIP: "HIP runtime error",
    IDP_E_NONPOS_DIST: "non-positive distance in a barrier term",
    IDP_E_NOT_SPD: "linear solve breakdown (Hessian not SPD)",
    IDP_E_CCD_ZERO: "zero time of impact",
    IDP_E_CAPACITY: "output capacity too small",
    IDP_E_INVALID_START: "degenerate tet at the start of the step",
    IDP_E_NO_GPU: "no HIP device",
}

EXPORTED = [
    "idp_ctx_create", "idp_ctx_destroy", "idp_last_error", "idp_device_count", "idp_fem_init", "idp_surface",
    "idp_initialize_oipc", "idp_set_mesh", "idp_set_dbc", "idp_set_dbc_targets", "idp_set_state", "idp_get_state", "idp_snapshot_state",
    "idp_restore_state", "idp_surface_sizes", "idp_advance_one_step", "idp_constraint_set", "idp_get_constraints",
    "idp_barrier_energy", "idp_barrier_gradient", "idp_barrier_hessian_local", "idp_min_dist2", "idp_ccd_stepsize",
    "idp_elastic_energy", "idp_elastic_gradient", "idp_elastic_hessian_local", "idp_newton_direction", "idp_get_bsr",
    "idp_profile_enable", "idp_profile_read", "idp_set_mesh_batch", "idp_batch_size", "idp_advance_batch",
    "idp_serve_configure", "idp_serve_start", "idp_serve_iterate", "idp_serve_result", "idp_set_launch_config",
    "idp_get_launch_config", "idp_set_accd_max_steps", "idp_set_accd_certificate", "idp_ccd_last_stats", "idp_barrier_energy_noalloc", "idp_get_surface", "idp_device_memory", "idp_set_barrier_team",
    "idp_get_barrier_team", "idp_load_velocity", "idp_shell_init", "idp_shell_surface", "idp_set_mesh_shell",
    "idp_normal_flow_force", "idp_set_normal_flow", "idp_newton_direction_flow", "idp_shell_hinges",
    "idp_set_shell_elasticity", "idp_set_friction", "idp_friction_basis", "idp_get_friction", "idp_friction_energy",
    "idp_friction_gradient", "idp_friction_hessian_local",
]

# PCG preconditioner for every solve of this process: 0 cluster block Jacobi
# with clusters matched over elastic and contact couplings, 1 plain 3x3 block
# Jacobi, 2 clusters matched over contact couplings only (IDP_PRECOND)
PRECOND = int(os.environ.get("IDP_PRECOND", "0"))
# 1: every PCG solve starts from zero instead of the previous direction (IDP_PCG_COLD)
PCG_COLD = int(os.environ.get("IDP_PCG_COLD", "0"))
# batched engine: ACCD steps per pair and Newton iteration before a hand-over
# (0 default, < 0 never; IDP_ACCD_BUDGET)
ACCD_BUDGET = int(os.environ.get("IDP_ACCD_BUDGET", "0"))
# batched engine: PCG hand-over (0 adaptive default, < 0 never; IDP_PCG_HANDOVER)
PCG_HANDOVER = int(os.environ.get("IDP_PCG_HANDOVER", "0"))
# coarse level of the one-mesh PCG preconditioner: free vertices per
# aggregate (0 library default, < 0 none; IDP_COARSE)
COARSE_VERTS = int(os.environ.get("IDP_COARSE", "0"))

# PCG launch configuration written by the autotuner (idp_amd/autotune.py):
# IDP_TUNING names the file, else tuning.json beside this module; absent or
# unreadable means the library defaults.  One configuration serves every
# context of the process, so an instance's result never depends on its mix.
TUNING_PATH = os.environ.get("IDP_TUNING", os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json"))
_tuned = None


def tuned_barrier_team(path=None):
    """Barrier Hessian team width from a tuning file, or None."""
    import json

    try:
        with open(path or TUNING_PATH) as f:
            t = int(json.load(f)["barrier_team"])
        return t if t in (3, 8, 16, 32, 64) else None
    except (OSError, ValueError, KeyError, TypeError):
        return None


def tuned_config(path=None):
    """(spmv_shape, upd_waves) from a tuning file, or None."""
    global _tuned
    import json

    if path is None and _tuned is not None:
        return _tuned or None
    p = path or TUNING_PATH
    cfg = ()
    try:
        with open(p) as f:
            d = json.load(f)
        sh, uw = int(d["spmv_shape"]), int(d["upd_waves"])
        if 0 <= sh <= 7 and 1 <= uw <= 8:
            cfg = (sh, uw)
    except (OSError, ValueError, KeyError, TypeError):
        cfg = ()
    if path is None:
        _tuned = cfg
    return cfg or None


K_FAMILIES = ["pcg", "barrier_hess", "tet_hess", "assembly", "constraint_set", "ccd", "energy", "gradient",
              "pcg_spmv", "pcg_update", "friction", "friction_basis"]


class IdpError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{_STATUS.get(code, 'error')} (status {code}){': ' + msg if msg else ''}")
        self.code = code


# ACCD step cap per pair and call (0: the library default, 2^20; < 0: none,
# as in the reference; IDP_ACCD_MAX_STEPS)
ACCD_MAX_STEPS = int(os.environ.get("IDP_ACCD_MAX_STEPS", "0"))
ACCD_MAX_STEPS_DEFAULT = 1 << 20


class SolverOpts(C.Structure):
    _fields_ = [("pcg_rtol", C.c_double), ("pcg_maxit", C.c_int), ("max_newton", C.c_int), ("verbose", C.c_int),
                ("precond", C.c_int), ("pcg_cold_start", C.c_int), ("accd_budget", C.c_int),
                ("pcg_handover", C.c_int), ("accd_max_steps", C.c_int), ("coarse_verts", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [
        ("newton_iters", C.c_int), ("line_search_trials", C.c_int), ("constraint_sets", C.c_int),
        ("gd_iters", C.c_int), ("pcg_iters", C.c_long), ("n_constraints", C.c_long), ("ccd_candidates", C.c_long),
        ("last_alpha", C.c_double), ("first_ccd_alpha", C.c_double), ("last_res", C.c_double),
        ("last_relres", C.c_double), ("min_dist2", C.c_double), ("E", C.c_double), ("kappa", C.c_double),
        ("ms_total", C.c_double), ("ms_constraint_set", C.c_double), ("ms_ccd", C.c_double),
        ("ms_energy", C.c_double), ("ms_gradient", C.c_double), ("ms_hessian", C.c_double),
        ("ms_solve", C.c_double), ("dbc_alpha", C.c_double), ("dbc_stiff", C.c_double),
        ("pcg_unconverged", C.c_int), ("accd_capped", C.c_long), ("n_friction", C.c_long),
        ("friction_updates", C.c_int),
        ("accd_certified", C.c_long),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * 16), ("launches", C.c_long * 16), ("work", C.c_double * 16),
                ("barrier_pairs", C.c_long * 8), ("barrier_clamped", C.c_long * 8), ("ccd_steps", C.c_double),
                ("ccd_max_steps", C.c_double)]

    def as_dict(self):
        out = {}
        for i, name in enumerate(K_FAMILIES):
            out[name] = dict(ms=self.ms[i], launches=self.launches[i], work=self.work[i])
        out["barrier_pairs"] = list(self.barrier_pairs)[:7]
        out["barrier_clamped"] = list(self.barrier_clamped)[:7]
        out["ccd_steps"] = self.ccd_steps
        out["ccd_max_steps"] = self.ccd_max_steps
        return out


_lib = None
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libidp_amd.so not built ({LIB_PATH}); run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.idp_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.idp_ctx_destroy.argtypes = [vp]
        L.idp_last_error.argtypes = [vp]
        L.idp_last_error.restype = C.c_char_p
        L.idp_device_count.argtypes = [ip]
        L.idp_fem_init.argtypes = [C.c_int, dp, C.c_int, ip, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp, dp]
        L.idp_surface.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, ip, ip, ip, ip]
        L.idp_initialize_oipc.argtypes = [C.c_double, C.c_int, dp, C.c_double, dp, dp]
        L.idp_set_mesh.argtypes = [vp, C.c_int, dp, C.c_int, ip, dp, dp, dp, dp, dp]
        L.idp_set_dbc.argtypes = [vp, C.c_int, ip]
        L.idp_set_dbc_targets.argtypes = [vp, C.c_int, ip, dp]
        L.idp_set_state.argtypes = [vp, dp, dp]
        L.idp_get_state.argtypes = [vp, dp, dp]
        L.idp_snapshot_state.argtypes = [vp]
        L.idp_restore_state.argtypes = [vp]
        L.idp_surface_sizes.argtypes = [vp, ip, ip, ip]
        L.idp_get_surface.argtypes = [vp, ip, ip, ip]
        L.idp_set_barrier_team.argtypes = [vp, C.c_int]
        L.idp_load_velocity.argtypes = [vp, dp, C.c_double]
        L.idp_get_barrier_team.argtypes = [vp, ip]
        L.idp_device_memory.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.idp_advance_one_step.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_double, dp, C.c_double, dp,
                                           C.c_int, C.POINTER(SolverOpts), C.POINTER(StepStats)]
        L.idp_set_accd_max_steps.argtypes = [vp, C.c_int]
        L.idp_set_accd_certificate.argtypes = [vp, C.c_int]
        L.idp_ccd_last_stats.argtypes = [vp, C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.idp_constraint_set.argtypes = [vp, dp, C.c_double, C.c_double, C.POINTER(C.c_long)]
        L.idp_get_constraints.argtypes = [vp, ip, C.c_long]
        L.idp_barrier_energy.argtypes = [vp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double, dp]
        L.idp_barrier_gradient.argtypes = [vp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double, dp]
        L.idp_barrier_energy_noalloc.argtypes = [vp, vp, C.c_long, vp, C.c_double, C.c_double, C.c_double, dp]
        L.idp_barrier_hessian_local.argtypes = [vp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double, C.c_int, dp,
                                                ip]
        L.idp_min_dist2.argtypes = [vp, dp, C.c_long, ip, C.c_double, dp, dp]
        L.idp_ccd_stepsize.argtypes = [vp, dp, dp, C.c_double, C.c_double, dp, C.POINTER(C.c_long)]
        L.idp_elastic_energy.argtypes = [vp, dp, C.c_double, dp, ip]
        L.idp_elastic_gradient.argtypes = [vp, dp, C.c_double, dp]
        L.idp_elastic_hessian_local.argtypes = [vp, dp, C.c_double, C.c_int, dp]
        L.idp_newton_direction.argtypes = [vp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                           C.POINTER(SolverOpts), dp, dp, dp, ip]
        L.idp_get_bsr.argtypes = [vp, C.POINTER(C.c_long), ip, ip, dp, C.c_long]
        L.idp_set_mesh_batch.argtypes = [vp, C.c_int, ip, ip, dp, ip, dp, dp, dp, dp, dp]
        L.idp_batch_size.argtypes = [vp, ip]
        L.idp_advance_batch.argtypes = [vp, C.c_double, dp, C.c_int, dp, dp, C.c_double, C.POINTER(SolverOpts), ip,
                                        C.POINTER(StepStats)]
        L.idp_serve_configure.argtypes = [vp, C.c_double, C.c_int, C.c_double, C.POINTER(SolverOpts)]
        L.idp_serve_start.argtypes = [vp, C.c_int, ip, dp, dp, dp, dp, dp]
        L.idp_serve_iterate.argtypes = [vp, ip, ip]
        L.idp_serve_result.argtypes = [vp, C.c_int, dp, dp, dp, C.POINTER(StepStats)]
        L.idp_profile_enable.argtypes = [vp, C.c_int]
        L.idp_set_launch_config.argtypes = [vp, C.c_int, C.c_int]
        L.idp_get_launch_config.argtypes = [vp, ip, ip]
        L.idp_profile_read.argtypes = [vp, C.POINTER(Profile)]
        L.idp_shell_init.argtypes = [C.c_int, dp, C.c_int, ip, C.c_double, C.c_double, C.c_double, C.c_double, ip,
                                     ip, dp, dp, dp, dp]
        L.idp_shell_surface.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, ip, ip, ip, ip, ip, ip, ip]
        L.idp_set_mesh_shell.argtypes = [vp, C.c_int, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip, dp, C.c_int, ip]
        L.idp_set_shell_elasticity.argtypes = [vp, dp, dp, dp, dp, C.c_double]
        L.idp_shell_hinges.argtypes = [C.c_int, dp, C.c_int, ip, ip, ip, dp, dp]
        L.idp_normal_flow_force.argtypes = [vp, dp, C.c_double, dp]
        L.idp_set_normal_flow.argtypes = [vp, C.c_double]
        L.idp_set_friction.argtypes = [vp, C.c_double, C.c_double, C.c_int, C.c_int, ip, dp]
        L.idp_friction_basis.argtypes = [vp, dp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double,
                                         C.POINTER(C.c_long)]
        L.idp_get_friction.argtypes = [vp, ip, dp, dp, dp, C.c_long]
        L.idp_friction_energy.argtypes = [vp, dp, C.c_double, dp]
        L.idp_friction_gradient.argtypes = [vp, dp, C.c_double, dp]
        L.idp_friction_hessian_local.argtypes = [vp, dp, C.c_double, C.c_int, dp]
        L.idp_newton_direction_flow.argtypes = L.idp_newton_direction.argtypes
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ip)


def _out_d(n):
    a = np.zeros(n, np.float64)
    return a, a.ctypes.data_as(dp)


def _out_i(n):
    a = np.zeros(n, np.int32)
    return a, a.ctypes.data_as(ip)


def device_count():
    n = C.c_int()
    lib().idp_device_count(C.byref(n))
    return n.value


def device_memory():
    """(bytes now, most bytes at once) held by the library's device buffers in this process."""
    a, b = C.c_longlong(), C.c_longlong()
    lib().idp_device_memory(C.byref(a), C.byref(b))
    return a.value, b.value


def fem_init(X, tets, density, E, nu):
    """Compute_Vol_And_Inv_Basis + Lame parameters + lumped mass (host)."""
    X, Xp = _d(X)
    tets, tp = _i(tets)
    nV, nT = X.size // 3, tets.size // 4
    vol, vp = _out_d(nT)
    IB, ibp = _out_d(9 * nT)
    lam, lp = _out_d(nT)
    mu, mp = _out_d(nT)
    mass, msp = _out_d(nV)
    rc = lib().idp_fem_init(nV, Xp, nT, tp, density, E, nu, vp, ibp, lp, mp, msp)
    if rc != IDP_OK:
        raise IdpError(rc)
    return vol, IB, lam, mu, mass


def surface(X, tets):
    X, Xp = _d(X)
    tets, tp = _i(tets)
    nV, nT = X.size // 3, tets.size // 4
    bn, bnp = _out_i(nV)
    be, bep = _out_i(12 * nT)
    bt, btp = _out_i(12 * nT)
    nb, ne, nf = C.c_int(), C.c_int(), C.c_int()
    rc = lib().idp_surface(nV, Xp, nT, tp, C.byref(nb), C.byref(ne), C.byref(nf), bnp, bep, btp)
    if rc != IDP_OK:
        raise IdpError(rc)
    return bn[: nb.value].copy(), be[: 2 * ne.value].reshape(-1, 2).copy(), bt[: 3 * nf.value].reshape(-1, 3).copy()


def shell_init(X, tris, density, E, nu, thickness, mass=None):
    """Initialize_Discrete_Shell (host): (kept triangles, mass, vol, lambda, mu);
    mass accumulates into a copy of the given one (zeros when None)."""
    X, Xp = _d(X)
    tris, tp = _i(tris)
    nV, nE = X.size // 3, tris.size // 3
    kept, kp = _out_i(3 * nE)
    m, mp = _d(np.zeros(nV) if mass is None else np.array(mass, dtype=np.float64))
    vol, vp = _out_d(nE)
    lam, lp = _out_d(nE)
    mu, mup = _out_d(nE)
    n = C.c_int()
    rc = lib().idp_shell_init(nV, Xp, nE, tp, density, E, nu, thickness, C.byref(n), kp, mp, vp, lp, mup)
    if rc != IDP_OK:
        raise IdpError(rc)
    k = n.value
    return kept[: 3 * k].reshape(-1, 3).copy(), m, vol[:k].copy(), lam[:k].copy(), mu[:k].copy()


def shell_surface(X, tris, tets=None):
    """Boundary (nodes, edges, triangles) of tets' surface + shell triangles (host)."""
    X, Xp = _d(X)
    tris, trp = _i(tris)
    tets, tp = _i(np.zeros((0, 4)) if tets is None else tets)
    nV, nT, nE = X.size // 3, tets.size // 4, tris.size // 3
    bn, bnp = _out_i(nV)
    be, bep = _out_i(2 * (3 * nE + 12 * nT))
    bt, btp = _out_i(3 * (nE + 4 * nT))
    nb, ne, nf = C.c_int(), C.c_int(), C.c_int()
    rc = lib().idp_shell_surface(nV, Xp, nT, tp, nE, trp, C.byref(nb), C.byref(ne), C.byref(nf), bnp, bep, btp)
    if rc != IDP_OK:
        raise IdpError(rc)
    return bn[: nb.value].copy(), be[: 2 * ne.value].reshape(-1, 2).copy(), bt[: 3 * nf.value].reshape(-1, 3).copy()


def shell_hinges(X, tris):
    """(hinges (nH, 4), rest (nH, 3), tri_IB (nE, 3)) of a shell (host)."""
    X, Xp = _d(X)
    tris, tp = _i(tris)
    nV, nE = X.size // 3, tris.size // 3
    cap = (3 * nE) // 2 + 1
    hg, hp = _out_i(4 * cap)
    hr, hrp = _out_d(3 * cap)
    ib, ibp = _out_d(3 * nE)
    n = C.c_int()
    rc = lib().idp_shell_hinges(nV, Xp, nE, tp, C.byref(n), hp, hrp, ibp)
    if rc != IDP_OK:
        raise IdpError(rc)
    k = n.value
    return hg[: 4 * k].reshape(-1, 4).copy(), hr[: 3 * k].reshape(-1, 3).copy(), ib.reshape(-1, 3)


def initialize_oipc(dhat, mass, stiff_mult=1.0):
    mass, mp = _d(mass)
    k, kp = _out_d(3)
    d2 = C.c_double()
    rc = lib().idp_initialize_oipc(dhat, mass.size, mp, stiff_mult, kp, C.byref(d2))
    if rc != IDP_OK:
        raise IdpError(rc)
    return d2.value, k


# Contexts still open when the interpreter exits are destroyed by an atexit
# hook, which runs before the interpreter tears its modules down and before
# the HIP runtime's own exit handlers; a context is never destroyed from a
# finalizer during interpreter shutdown.
_live = weakref.WeakSet()


@atexit.register
def _close_all():
    for ctx in list(_live):
        ctx.close()


class Context:
    """One device context: mesh, state and work buffers on one HIP stream.
    Close it with close() (or a with-block); contexts left open are closed at
    interpreter exit."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().idp_ctx_create(device, C.byref(self._h))
        if rc != IDP_OK:
            raise IdpError(rc)
        _live.add(self)
        self.nV = 0
        cfg = tuned_config()
        if cfg and "IDP_SPMV_SHAPE" not in os.environ:  # an explicit shape (tuning runs) wins
            self.set_launch_config(*cfg)
        team = tuned_barrier_team()
        if team:
            self.set_barrier_team(team)

    def set_barrier_team(self, lanes):
        """Lanes per constraint in the barrier Hessian's projection stage (8, 16, 32, 64)."""
        self._check(lib().idp_set_barrier_team(self._h, int(lanes)))

    def barrier_team(self):
        t = C.c_int()
        self._check(lib().idp_get_barrier_team(self._h, C.byref(t)))
        return t.value

    def set_launch_config(self, spmv_shape, upd_waves):
        """PCG launch configuration (see idp_set_launch_config)."""
        self._check(lib().idp_set_launch_config(self._h, int(spmv_shape), int(upd_waves)))

    def launch_config(self):
        sh, uw = C.c_int(), C.c_int()
        self._check(lib().idp_get_launch_config(self._h, C.byref(sh), C.byref(uw)))
        return sh.value, uw.value

    def close(self):
        if self._h:
            h, self._h = self._h, C.c_void_p()
            _live.discard(self)
            lib().idp_ctx_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        # during interpreter shutdown the runtime may already be gone: the
        # atexit hook has closed every context that was still reachable
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise IdpError(rc, lib().idp_last_error(self._h).decode())
        return rc

    def set_mesh(self, X_rest, tets, IB, vol, lam, mu, mass):
        X, Xp = _d(X_rest)
        t, tp = _i(tets)
        ib, ibp = _d(IB)
        v, vp = _d(vol)
        l, lp = _d(lam)
        m, mp = _d(mu)
        ms, msp = _d(mass)
        self.nV = X.size // 3
        self._check(lib().idp_set_mesh(self._h, self.nV, Xp, t.size // 4, tp, ibp, vp, lp, mp, msp))

    def set_mesh_shell(self, X_rest, mass, tris, tri_vol, tets=None, IB=None, vol=None, lam=None, mu=None,
                       hinges=None):
        """A mesh of shell triangles (and optionally tets and hinges): idp_set_mesh_shell."""
        X, Xp = _d(X_rest)
        m, mp = _d(mass)
        tr, trp = _i(tris)
        tv, tvp = _d(tri_vol)
        hg, hgp = _i(np.zeros((0, 4)) if hinges is None else hinges)
        nT = 0 if tets is None else np.asarray(tets).size // 4
        args = [None] * 5
        if nT:  # the arrays stay referenced by these locals until the call returns
            te, tep = _i(tets)
            ib, ibp = _d(IB)
            vo, vop = _d(vol)
            la, lap = _d(lam)
            mu_, mup = _d(mu)
            args = [tep, ibp, vop, lap, mup]
        self._check(lib().idp_set_mesh_shell(self._h, X.size // 3, Xp, mp, nT, *args, tr.size // 3, trp, tvp,
                                             hg.size // 4, hgp))
        self.nV = X.size // 3
        self.nT = nT

    def set_shell_elasticity(self, tri_IB, tri_lambda, tri_mu, hinge_rest, bend_k):
        ib, ibp = _d(tri_IB)
        la, lap = _d(tri_lambda)
        mu_, mup = _d(tri_mu)
        hr, hrp = _d(np.zeros(0) if hinge_rest is None else hinge_rest)
        self._check(lib().idp_set_shell_elasticity(self._h, ibp, lap, mup, hrp, float(bend_k)))

    def normal_flow_force(self, magnitude, X=None):
        """Update_Normal_Flow_Neumann on the device at X (None: current x)."""
        b, bp = _out_d(3 * self.nV)
        xp = None
        if X is not None:
            X, xp = _d(X)
        self._check(lib().idp_normal_flow_force(self._h, xp, magnitude, bp))
        return b.reshape(-1, 3)

    def set_normal_flow(self, magnitude):
        self._check(lib().idp_set_normal_flow(self._h, magnitude))

    def set_friction(self, mu, epsv2=1e-6, fric_iters=1, comp_range=None, mu_comp=None):
        """Lagged friction of this context's steps (mu, epsv2, fricIterAmt,
        compNodeRange, muComp of the step); mu 0 without components: off."""
        if comp_range is None:
            self._check(lib().idp_set_friction(self._h, mu, epsv2, int(fric_iters), 0, None, None))
            return
        cr, crp = _i(np.asarray(comp_range))
        mc, mcp = _d(np.asarray(mu_comp, dtype=np.float64).ravel())
        if mc.size != cr.size * cr.size:
            raise ValueError("mu_comp must hold ncomp^2 coefficients")
        self._check(lib().idp_set_friction(self._h, mu, epsv2, int(fric_iters), cr.size, crp, mcp))

    def friction_basis(self, X, Xn, cs, dHat2, kappa, thickness=0.0):
        """Compute_Friction_Basis at X on the constraints cs; Xn is kept as the
        start positions of the friction stage calls.  Returns the count."""
        X, Xp = _d(X)
        xn, xnp = _d(Xn)
        cs, csp = _i(cs)
        nf = C.c_long()
        self._check(lib().idp_friction_basis(self._h, Xp, xnp, cs.size // 4, csp, dHat2, kappa, thickness,
                                             C.byref(nf)))
        return nf.value

    def get_friction(self, nf):
        """(fcs (nf, 4), closest points (nf, 2), bases (nf, 6): column 0 then 1, normal forces)."""
        fcs, fp = _out_i(4 * max(nf, 1))
        cp, cpp = _out_d(2 * max(nf, 1))
        b, bp = _out_d(6 * max(nf, 1))
        N, Np = _out_d(max(nf, 1))
        self._check(lib().idp_get_friction(self._h, fp, cpp, bp, Np, nf))
        return fcs[: 4 * nf].reshape(nf, 4), cp[: 2 * nf].reshape(nf, 2), b[: 6 * nf].reshape(nf, 6), N[:nf]

    def friction_energy(self, X, h):
        X, Xp = _d(X)
        E, Ep = _out_d(1)
        self._check(lib().idp_friction_energy(self._h, Xp, h, Ep))
        return float(E[0])

    def friction_gradient(self, X, h):
        X, Xp = _d(X)
        g, gp = _out_d(X.size)
        self._check(lib().idp_friction_gradient(self._h, Xp, h, gp))
        return g

    def friction_hessian_local(self, X, h, nf, project=True):
        X, Xp = _d(X)
        H, Hp = _out_d(144 * max(nf, 1))
        self._check(lib().idp_friction_hessian_local(self._h, Xp, h, int(project), Hp))
        return H[: 144 * nf].reshape(nf, 144)

    def set_mesh_batch(self, meshes):
        """meshes: list of (X_rest (V,3), tets (T,4) local ids, IB, vol, lam, mu, mass)."""
        nV = np.array([np.asarray(m[0]).reshape(-1, 3).shape[0] for m in meshes], np.int32)
        nT = np.array([np.asarray(m[1]).reshape(-1, 4).shape[0] for m in meshes], np.int32)
        cat = [np.ascontiguousarray(np.concatenate([np.asarray(m[k], dtype=np.float64 if k != 1 else np.int32)
                                                    .ravel() for m in meshes])) for k in range(7)]
        X, Xp = _d(cat[0])
        t, tp = _i(cat[1])
        ib, ibp = _d(cat[2])
        v, vp = _d(cat[3])
        l, lp = _d(cat[4])
        m, mp = _d(cat[5])
        ms, msp = _d(cat[6])
        self.nV = int(nV.sum())
        self.batch_nV = nV
        self._check(lib().idp_set_mesh_batch(self._h, len(meshes), nV.ctypes.data_as(ip), nT.ctypes.data_as(ip), Xp,
                                             tp, ibp, vp, lp, mp, msp))

    def advance_batch(self, h, tol, dHat2, kappa3, thickness=0.0, with_collision=True, pcg_rtol=1e-8,
                      pcg_maxit=10000, max_newton=0, verbose=0, accd_budget=None, pcg_handover=None,
                      accd_max_steps=None):
        """One step of every instance; kappa3 (B,3) is updated in place.
        verbose > 0 prints a progress line to stderr every `verbose` Newton iterations;
        accd_budget: ACCD steps per pair and Newton iteration before a hand-over
        (None: the process default, 0: the library default, < 0: never);
        pcg_handover: None the process default, 0 adaptive, < 0 never.
        Returns (newton iterations (B,), list of per-instance stats dicts)."""
        B = len(self.batch_nV)
        tol = np.ascontiguousarray(tol, dtype=np.float64)
        d2 = np.ascontiguousarray(dHat2, dtype=np.float64)
        if not (kappa3.flags["C_CONTIGUOUS"] and kappa3.dtype == np.float64 and kappa3.size == 3 * B):
            raise ValueError("kappa3 must be a contiguous float64 (B,3) array")
        its = np.zeros(B, np.int32)
        st = (StepStats * B)()
        opts = SolverOpts(pcg_rtol, pcg_maxit, max_newton, verbose, PRECOND, PCG_COLD,
                          ACCD_BUDGET if accd_budget is None else accd_budget,
                          PCG_HANDOVER if pcg_handover is None else pcg_handover,
                          ACCD_MAX_STEPS if accd_max_steps is None else accd_max_steps)
        self._check(lib().idp_advance_batch(self._h, h, tol.ctypes.data_as(dp), int(with_collision),
                                            d2.ctypes.data_as(dp), kappa3.ctypes.data_as(dp), thickness,
                                            C.byref(opts), its.ctypes.data_as(ip), st))
        return its, [x.as_dict() for x in st]

    # ---- continuous batching over the slots of a batch mesh ----------------
    def serve_configure(self, h, with_collision=True, thickness=0.0, pcg_rtol=1e-8, pcg_maxit=10000, max_newton=0,
                        accd_budget=None, pcg_handover=None, accd_max_steps=None):
        opts = SolverOpts(pcg_rtol, pcg_maxit, max_newton, 0, PRECOND, PCG_COLD,
                          ACCD_BUDGET if accd_budget is None else accd_budget,
                          PCG_HANDOVER if pcg_handover is None else pcg_handover,
                          ACCD_MAX_STEPS if accd_max_steps is None else accd_max_steps)
        self._check(lib().idp_serve_configure(self._h, h, int(with_collision), thickness, C.byref(opts)))

    def serve_start(self, slots, X, v, tol, dHat2, kappa3):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        Xa, Xp = _d(np.concatenate([np.asarray(x).ravel() for x in X]))
        va, vp = _d(np.concatenate([np.asarray(x).ravel() for x in v]))
        t = np.ascontiguousarray(tol, dtype=np.float64)
        d2 = np.ascontiguousarray(dHat2, dtype=np.float64)
        k3 = np.ascontiguousarray(kappa3, dtype=np.float64).ravel()
        self._check(lib().idp_serve_start(self._h, slots.size, slots.ctypes.data_as(ip), Xp, vp,
                                          t.ctypes.data_as(dp), d2.ctypes.data_as(dp), k3.ctypes.data_as(dp)))

    def serve_iterate(self):
        """One Newton iteration of all busy slots: (busy count, finished mask)."""
        B = len(self.batch_nV)
        fin = np.zeros(B, np.int32)
        n = C.c_int()
        self._check(lib().idp_serve_iterate(self._h, C.byref(n), fin.ctypes.data_as(ip)))
        return n.value, fin.astype(bool)

    def serve_result(self, slot):
        nv = int(self.batch_nV[slot])
        X, Xp = _out_d(3 * nv)
        v, vp = _out_d(3 * nv)
        k3, kp = _out_d(3)
        st = StepStats()
        self._check(lib().idp_serve_result(self._h, slot, Xp, vp, kp, C.byref(st)))
        return X.reshape(-1, 3), v.reshape(-1, 3), k3, st.as_dict()

    def set_dbc(self, idx):
        a, ap = _i(idx)
        self._check(lib().idp_set_dbc(self._h, a.size, ap))

    def set_dbc_targets(self, idx, targets):
        """Moving Dirichlet nodes: idx (n,) driven to targets (n, 3) in every step."""
        a, ap = _i(idx)
        t, tp = _d(np.asarray(targets, dtype=np.float64).reshape(-1))
        if t.size != 3 * a.size:
            raise ValueError("targets must be (n, 3)")
        self._check(lib().idp_set_dbc_targets(self._h, a.size, ap, tp))

    def surface_sizes(self):
        nb, ne, nf = C.c_int(), C.c_int(), C.c_int()
        self._check(lib().idp_surface_sizes(self._h, C.byref(nb), C.byref(ne), C.byref(nf)))
        return nb.value, ne.value, nf.value

    def get_surface(self):
        """(boundary nodes, edges (ne, 2), triangles (nf, 3)) extracted by set_mesh."""
        nb, ne, nf = self.surface_sizes()
        bn, bnp = _out_i(max(nb, 1))
        be, bep = _out_i(max(2 * ne, 1))
        bt, btp = _out_i(max(3 * nf, 1))
        self._check(lib().idp_get_surface(self._h, bnp, bep, btp))
        return bn[:nb].copy(), be[: 2 * ne].reshape(-1, 2).copy(), bt[: 3 * nf].reshape(-1, 3).copy()

    def set_state(self, X=None, v=None):
        Xa, Xp = _d(X) if X is not None else (None, None)
        va, vp = _d(v) if v is not None else (None, None)
        self._check(lib().idp_set_state(self._h, Xp, vp))

    def load_velocity(self, X_target, h):
        """v = (X_target - x) / h on the device (Load_Velocity_X0)."""
        t, tp = _d(X_target)
        self._check(lib().idp_load_velocity(self._h, tp, h))

    def get_state(self):
        X, Xp = _out_d(3 * self.nV)
        v, vp = _out_d(3 * self.nV)
        self._check(lib().idp_get_state(self._h, Xp, vp))
        return X.reshape(-1, 3), v.reshape(-1, 3)

    def profile(self, on=True):
        self._check(lib().idp_profile_enable(self._h, int(on)))

    def profile_read(self):
        p = Profile()
        self._check(lib().idp_profile_read(self._h, C.byref(p)))
        return p.as_dict()

    def snapshot(self):
        self._check(lib().idp_snapshot_state(self._h))

    def restore(self):
        self._check(lib().idp_restore_state(self._h))

    def advance(self, h, tol, dHat2, kappa, thickness=0.0, with_collision=True, pcg_rtol=1e-8, pcg_maxit=10000,
                max_newton=0, verbose=0, body_force=None, flow=False, accd_max_steps=None, coarse_verts=None):
        """Advance_One_Step_IE_Discrete_Shell (flow=False: the Hinge binding,
        True: the Flow binding); kappa (3,) is updated in place; body_force
        (nV, 3) or None."""
        kappa_arr = np.asarray(kappa, dtype=np.float64)
        if not (kappa_arr.flags["C_CONTIGUOUS"] and kappa_arr.dtype == np.float64 and kappa_arr.size == 3):
            raise ValueError("kappa must be a contiguous float64 array of 3")
        bf, bfp = (None, None) if body_force is None else _d(np.asarray(body_force, dtype=np.float64).ravel())
        if bf is not None and bf.size != 3 * self.nV:
            raise ValueError("body_force must be (nV, 3)")
        opts = SolverOpts(pcg_rtol, pcg_maxit, max_newton, verbose, PRECOND, PCG_COLD, ACCD_BUDGET, PCG_HANDOVER,
                          ACCD_MAX_STEPS if accd_max_steps is None else accd_max_steps,
                          COARSE_VERTS if coarse_verts is None else coarse_verts)
        st = StepStats()
        it = self._check(lib().idp_advance_one_step(self._h, h, tol, int(with_collision), dHat2,
                                                    kappa_arr.ctypes.data_as(dp), thickness, bfp, int(bool(flow)),
                                                    C.byref(opts), C.byref(st)))
        return it, st.as_dict()

    def set_accd_max_steps(self, n):
        """ACCD step cap of the stage calls (0 default, < 0 none)."""
        self._check(lib().idp_set_accd_max_steps(self._h, int(n)))

    def set_accd_certificate(self, on):
        """No-hit certificate of long ACCD pairs (on by default)."""
        self._check(lib().idp_set_accd_certificate(self._h, int(bool(on))))

    def ccd_last_stats(self):
        """(capped, certified) ACCD pairs of the last CCD call."""
        a, b = C.c_long(), C.c_long()
        self._check(lib().idp_ccd_last_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- stage entry points -------------------------------------------------
    def constraint_set(self, X, dHat2, thickness=0.0):
        X, Xp = _d(X)
        n = C.c_long()
        self._check(lib().idp_constraint_set(self._h, Xp, dHat2, thickness, C.byref(n)))
        cs, csp = _out_i(4 * max(n.value, 1))
        self._check(lib().idp_get_constraints(self._h, csp, n.value))
        return cs[: 4 * n.value].reshape(-1, 4)

    def barrier_energy(self, X, cs, dHat2, kappa, thickness=0.0):
        X, Xp = _d(X)
        cs, csp = _i(cs)
        E = C.c_double()
        self._check(lib().idp_barrier_energy(self._h, Xp, cs.size // 4, csp, dHat2, kappa, thickness, C.byref(E)))
        return E.value

    def barrier_energy_noalloc(self, X_dev, cs_dev, n, dHat2, kappa, thickness=0.0):
        """idp_barrier_energy_noalloc: X_dev and cs_dev are device addresses
        (ints) of 3*nV doubles and 4*n ints."""
        E = C.c_double()
        self._check(lib().idp_barrier_energy_noalloc(self._h, C.c_void_p(X_dev), int(n), C.c_void_p(cs_dev), dHat2,
                                                     kappa, thickness, C.byref(E)))
        return E.value

    def barrier_gradient(self, X, cs, dHat2, kappa, thickness=0.0):
        X, Xp = _d(X)
        cs, csp = _i(cs)
        g, gp = _out_d(X.size)
        self._check(lib().idp_barrier_gradient(self._h, Xp, cs.size // 4, csp, dHat2, kappa, thickness, gp))
        return g

    def barrier_hessian_local(self, X, cs, dHat2, kappa, thickness=0.0, project=True):
        X, Xp = _d(X)
        cs, csp = _i(cs)
        n = cs.size // 4
        H, Hp = _out_d(144 * max(n, 1))
        vid, vp = _out_i(4 * max(n, 1))
        self._check(lib().idp_barrier_hessian_local(self._h, Xp, n, csp, dHat2, kappa, thickness, int(project), Hp,
                                                    vp))
        return H[: 144 * n].reshape(n, 144), vid[: 4 * n].reshape(n, 4)

    def min_dist2(self, X, cs, thickness=0.0):
        X, Xp = _d(X)
        cs, csp = _i(cs)
        n = cs.size // 4
        d, dpp = _out_d(max(n, 1))
        m = C.c_double()
        self._check(lib().idp_min_dist2(self._h, Xp, n, csp, thickness, dpp, C.byref(m)))
        return d[:n], m.value

    def ccd(self, X, p, alpha, thickness=0.0):
        X, Xp = _d(X)
        p, pp = _d(p)
        out = C.c_double()
        nc = C.c_long()
        self._check(lib().idp_ccd_stepsize(self._h, Xp, pp, thickness, alpha, C.byref(out), C.byref(nc)))
        return out.value, nc.value

    def elastic_energy(self, X, w):
        X, Xp = _d(X)
        E = C.c_double()
        ok = C.c_int()
        self._check(lib().idp_elastic_energy(self._h, Xp, w, C.byref(E), C.byref(ok)))
        return E.value if ok.value else None

    def elastic_gradient(self, X, w):
        X, Xp = _d(X)
        g, gp = _out_d(X.size)
        self._check(lib().idp_elastic_gradient(self._h, Xp, w, gp))
        return g

    def elastic_hessian_local(self, X, w, nT, project=True):
        X, Xp = _d(X)
        H, Hp = _out_d(144 * nT)
        self._check(lib().idp_elastic_hessian_local(self._h, Xp, w, int(project), Hp))
        return H.reshape(nT, 144)

    def newton_direction(self, X, xtilde, h, dHat2, kappa, thickness=0.0, pcg_rtol=1e-10, pcg_maxit=20000,
                         flow=False, coarse_verts=None):
        """(g, direction, relative residual, PCG iterations) of the Newton system
        at X; flow: the flow variant's terms (idp_newton_direction_flow)."""
        X, Xp = _d(X)
        xt, xtp = _d(xtilde)
        g, gp = _out_d(X.size)
        sol, sp = _out_d(X.size)
        rel = C.c_double()
        it = C.c_int()
        opts = SolverOpts(pcg_rtol, pcg_maxit, 0, 0, PRECOND, PCG_COLD, ACCD_BUDGET, PCG_HANDOVER, ACCD_MAX_STEPS,
                          COARSE_VERTS if coarse_verts is None else coarse_verts)
        fn = lib().idp_newton_direction_flow if flow else lib().idp_newton_direction
        self._check(fn(self._h, Xp, xtp, h, dHat2, kappa, thickness, C.byref(opts), gp, sp, C.byref(rel),
                       C.byref(it)))
        return g, sol, rel.value, it.value

    def get_bsr(self):
        nnzb = C.c_long()
        rc = lib().idp_get_bsr(self._h, C.byref(nnzb), None, None, None, 0)
        if rc not in (IDP_OK, IDP_E_CAPACITY):
            self._check(rc)
        rp, rpp = _out_i(self.nV + 1)
        col, cp = _out_i(max(nnzb.value, 1))
        val, vp = _out_d(9 * max(nnzb.value, 1))
        self._check(lib().idp_get_bsr(self._h, C.byref(nnzb), rpp, cp, vp, nnzb.value))
        return rp, col[: nnzb.value], val[: 9 * nnzb.value].reshape(-1, 3, 3)
