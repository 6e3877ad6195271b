"""ctypes loader for the oracle (test infrastructure only).

Used by tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
bench.py to check or time the CPU restatement; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class Stats(C.Structure):
    _fields_ = [
        ("n_newton", C.c_int),
        ("line_search_trials", C.c_int),
        ("n_constraint_sets", C.c_int),
        ("linear_iters", C.c_long),
        ("n_constraints", C.c_long),
        ("last_alpha", C.c_double),
        ("first_ccd_alpha", C.c_double),
        ("last_res", C.c_double),
        ("last_relres", C.c_double),
        ("min_dist2", C.c_double),
        ("E", C.c_double),
        ("t_constraint_set", C.c_double),
        ("t_ccd", C.c_double),
        ("t_solve", C.c_double),
        ("t_hessian", C.c_double),
        ("t_gradient", C.c_double),
        ("t_energy", C.c_double),
        ("dbc_alpha", C.c_double),
        ("dbc_stiff", C.c_double),
        ("n_friction", C.c_long),
        ("friction_updates", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_surface.argtypes = [C.c_int, dp, C.c_int, ip, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), ip, ip, ip]
        L.orc_constraint_set.restype = C.c_long
        L.orc_constraint_set.argtypes = [C.c_int, dp, dp, u8p, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_double,
                                         C.c_double, C.c_int, ip, C.c_long]
        L.orc_barrier_energy.argtypes = [dp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double,
                                         C.POINTER(C.c_double), C.c_void_p]
        L.orc_barrier_gradient.argtypes = [dp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double, C.c_int, dp]
        L.orc_barrier_hessian_local.argtypes = [dp, dp, C.c_long, ip, C.c_double, C.c_double, C.c_double,
                                                C.c_int, dp, ip]
        L.orc_min_dist2.argtypes = [dp, C.c_long, ip, C.c_double, C.c_void_p, C.POINTER(C.c_double)]
        L.orc_ccd_stepsize.argtypes = [C.c_int, dp, dp, u8p, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_double,
                                       C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_long)]
        L.orc_geom_probe.argtypes = [C.c_int, dp, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_fem_init.argtypes = [C.c_int, dp, C.c_int, ip, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp,
                                   dp]
        L.orc_elastic_energy.argtypes = [dp, C.c_int, ip, dp, dp, dp, dp, C.c_double, C.POINTER(C.c_double)]
        L.orc_elastic_gradient.argtypes = [dp, C.c_int, ip, dp, dp, dp, dp, C.c_double, dp]
        L.orc_elastic_gradient.restype = None
        L.orc_elastic_hessian_local.argtypes = [dp, C.c_int, ip, dp, dp, dp, dp, C.c_double, C.c_int, dp]
        L.orc_elastic_hessian_local.restype = None
        L.orc_advance_one_step.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip,
                                           C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_void_p, C.c_double, C.c_double,
                                           C.c_int, C.c_double, dp, C.c_double, C.c_double, C.c_int,
                                           C.c_void_p, C.c_int, C.POINTER(Stats)]
        L.orc_set_accd_max_steps.argtypes = [C.c_int]
        L.orc_set_accd_max_steps.restype = None
        L.orc_newton_system.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip, C.c_int,
                                        ip, C.c_int, ip, C.c_int, ip, C.c_double, C.c_double, C.c_double,
                                        C.c_double, ip, ip, dp, dp, C.c_long]
        L.orc_newton_system.restype = C.c_long
        L.orc_shell_init.argtypes = [C.c_int, dp, C.c_int, ip, C.c_double, C.c_double, C.c_double, C.c_double, ip,
                                     dp, dp, dp, dp]
        L.orc_shell_surface.argtypes = [C.c_int, dp, C.c_int, ip, C.c_int, ip, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int), ip, ip, ip]
        L.orc_normal_flow_force.argtypes = [C.c_int, dp, C.c_int, ip, dp, C.c_double, dp]
        L.orc_normal_flow_force.restype = None
        L.orc_flow_terms.argtypes = [C.c_int, dp, C.c_int, ip, dp, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_flow_terms.restype = None
        L.orc_advance_shell_flow.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, C.c_int, ip, C.c_int, ip,
                                             C.c_int, ip, C.c_int, ip, C.c_double, C.c_double, C.c_int, C.c_double,
                                             dp, C.c_double, C.c_double, C.c_int, C.c_void_p, C.POINTER(Stats)]
        L.orc_newton_system_shell_flow.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, C.c_int, ip, C.c_int,
                                                   ip, C.c_int, ip, C.c_int, ip, C.c_double, C.c_double, C.c_double,
                                                   C.c_double, ip, ip, dp, dp, C.c_long]
        L.orc_newton_system_shell_flow.restype = C.c_long
        L.orc_shell_rest_fff.argtypes = [dp, C.c_int, ip, dp]
        L.orc_shell_rest_fff.restype = None
        L.orc_shell_hinges.argtypes = [dp, C.c_int, ip, C.c_int, ip, ip, dp]
        L.orc_shell_elastic_terms.argtypes = [C.c_int, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip, dp, C.c_double,
                                              C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p]
        L.orc_shell_elastic_terms.restype = None
        L.orc_dihedral.argtypes = [dp, C.c_void_p]
        L.orc_dihedral.restype = C.c_double
        L.orc_advance_shell.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip, dp,
                                        C.c_double, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_double,
                                        C.c_double, C.c_int, C.c_double, dp, C.c_double, C.c_double, C.c_int,
                                        C.c_void_p, C.POINTER(Stats)]
        L.orc_newton_system_shell.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip, dp,
                                              C.c_double, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_int, ip,
                                              C.c_double, C.c_double, C.c_double, C.c_double, ip, ip, dp, dp,
                                              C.c_long]
        L.orc_newton_system_shell.restype = C.c_long
        L.orc_advance_shell_friction.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip,
                                                 dp, C.c_double, C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_int, ip,
                                                 C.c_double, C.c_double, C.c_double, dp, C.c_double, C.c_double,
                                                 C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int,
                                                 C.POINTER(Stats)]
        L.orc_friction_basis.argtypes = [dp, C.c_long, ip, C.c_double, C.c_double, C.c_double, ip, dp, dp, dp]
        L.orc_friction_basis.restype = C.c_long
        L.orc_friction_coef.argtypes = [C.c_long, ip, C.c_int, ip, dp, dp]
        L.orc_friction_energy.argtypes = [dp, dp, C.c_long, ip, dp, dp, dp, C.c_double, C.c_double]
        L.orc_friction_energy.restype = C.c_double
        L.orc_friction_gradient.argtypes = [dp, dp, C.c_long, ip, dp, dp, dp, C.c_double, C.c_double, dp]
        L.orc_friction_gradient.restype = None
        L.orc_friction_hessian_local.argtypes = [dp, dp, C.c_long, ip, dp, dp, dp, C.c_double, C.c_double, C.c_int,
                                                 dp, ip]
        L.orc_friction_hessian_local.restype = None
        L.orc_advance_friction.argtypes = [C.c_int, dp, dp, dp, dp, C.c_int, ip, dp, dp, dp, dp, C.c_int, ip,
                                           C.c_int, ip, C.c_int, ip, C.c_int, ip, C.c_void_p, C.c_double, C.c_double,
                                           C.c_int, C.c_double, dp, C.c_double, C.c_double, C.c_int,
                                           C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p, C.POINTER(Stats)]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64).ravel()


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ravel()


def surface(X, tets):
    X = _f(X)
    tets = _i(tets)
    nV, nT = X.size // 3, tets.size // 4
    bn = np.zeros(nV, np.int32)
    bt = np.zeros(3 * 4 * nT, np.int32)
    be = np.zeros(2 * 6 * nT, np.int32)
    nb, ne, nf = C.c_int(), C.c_int(), C.c_int()
    lib().orc_surface(nV, X, nT, tets, C.byref(nb), C.byref(ne), C.byref(nf), bn, be, bt)
    return bn[: nb.value].copy(), be[: 2 * ne.value].reshape(-1, 2).copy(), bt[: 3 * nf.value].reshape(-1, 3).copy()


class Mesh:
    """Host-side bundle of what the reference keeps in its Storage objects."""

    def __init__(self, X, tets, density, E, nu):
        self.X = _f(X).copy()
        self.X0 = self.X.copy()
        self.tets = _i(tets)
        self.nV, self.nT = self.X.size // 3, self.tets.size // 4
        self.vol = np.zeros(self.nT)
        self.IB = np.zeros(9 * self.nT)
        self.lam = np.zeros(self.nT)
        self.mu = np.zeros(self.nT)
        self.mass = np.zeros(self.nV)
        rc = lib().orc_fem_init(self.nV, self.X, self.nT, self.tets, density, E, nu, self.vol, self.IB, self.lam,
                                self.mu, self.mass)
        if rc != 0:
            raise RuntimeError(f"orc_fem_init failed: {rc}")
        self.bnode, self.bedge, self.btri = surface(self.X, self.tets)
        self.dbc = np.zeros(self.nV, np.uint8)

    def constraint_set(self, X, dHat2, thickness=0.0, use_hash=True):
        X = _f(X)
        cap = 1 << 16
        while True:
            out = np.zeros(4 * cap, np.int32)
            n = lib().orc_constraint_set(self.nV, X, self.X0, self.dbc, self.bnode.size, self.bnode,
                                         self.bedge.shape[0], _i(self.bedge), self.btri.shape[0], _i(self.btri),
                                         dHat2, thickness, int(use_hash), out, cap)
            if n >= 0:
                return out[: 4 * n].reshape(-1, 4).copy()
            cap = -n

    def barrier_energy(self, X, cs, dHat2, kappa, thickness=0.0):
        E = C.c_double()
        rc = lib().orc_barrier_energy(_f(X), self.X0, cs.shape[0], _i(cs), dHat2, kappa, thickness, C.byref(E), None)
        if rc:
            raise RuntimeError("non-positive distance")
        return E.value

    def barrier_gradient(self, X, cs, dHat2, kappa, thickness=0.0):
        g = np.zeros(3 * self.nV)
        rc = lib().orc_barrier_gradient(_f(X), self.X0, cs.shape[0], _i(cs), dHat2, kappa, thickness, self.nV, g)
        if rc:
            raise RuntimeError("non-positive distance")
        return g

    def barrier_hessian_local(self, X, cs, dHat2, kappa, thickness=0.0, project=True):
        n = cs.shape[0]
        H = np.zeros(144 * n)
        vid = np.zeros(4 * n, np.int32)
        rc = lib().orc_barrier_hessian_local(_f(X), self.X0, n, _i(cs), dHat2, kappa, thickness, int(project), H,
                                             vid)
        if rc:
            raise RuntimeError("non-positive distance")
        return H.reshape(n, 144), vid.reshape(n, 4)

    def min_dist2(self, X, cs, thickness=0.0):
        d = np.zeros(cs.shape[0])
        m = C.c_double()
        lib().orc_min_dist2(_f(X), cs.shape[0], _i(cs), thickness, d.ctypes.data_as(C.c_void_p), C.byref(m))
        return d, m.value

    def ccd(self, X, p, alpha, thickness=0.0):
        out = C.c_double()
        nc = C.c_long()
        lib().orc_ccd_stepsize(self.nV, _f(X), _f(p), self.dbc, self.bnode.size, self.bnode, self.bedge.shape[0],
                               _i(self.bedge), self.btri.shape[0], _i(self.btri), thickness, alpha, C.byref(out),
                               C.byref(nc))
        return out.value, nc.value

    def elastic_energy(self, X, w):
        E = C.c_double()
        ok = lib().orc_elastic_energy(_f(X), self.nT, self.tets, self.IB, self.vol, self.lam, self.mu, w, C.byref(E))
        return (E.value if ok else None)

    def elastic_gradient(self, X, w):
        g = np.zeros(3 * self.nV)
        lib().orc_elastic_gradient(_f(X), self.nT, self.tets, self.IB, self.vol, self.lam, self.mu, w, g)
        return g

    def elastic_hessian_local(self, X, w, project=True):
        H = np.zeros(144 * self.nT)
        lib().orc_elastic_hessian_local(_f(X), self.nT, self.tets, self.IB, self.vol, self.lam, self.mu, w,
                                        int(project), H)
        return H.reshape(self.nT, 144)

    def newton_system(self, X, Xtilde, h, dHat2, kappa, thickness=0.0, dbc_idx=None):
        """(rowptr, col, val (nnzb,3,3), g) of the projected Newton system at X."""
        X = _f(X)
        Xt = _f(Xtilde)
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        cap = 64 * self.nV + 1024
        while True:
            rp = np.zeros(self.nV + 1, np.int32)
            col = np.zeros(cap, np.int32)
            val = np.zeros(9 * cap)
            g = np.zeros(3 * self.nV)
            n = lib().orc_newton_system(self.nV, X, Xt, self.X0, self.mass, self.nT, self.tets, self.IB, self.vol,
                                        self.lam, self.mu, self.bnode.size, self.bnode, self.bedge.shape[0],
                                        _i(self.bedge), self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx,
                                        h, dHat2, kappa, thickness, rp, col, val, g, cap)
            if n >= 0:
                return rp, col[:n], val[: 9 * n].reshape(n, 3, 3), g
            cap = -n

    def advance_friction(self, X, vel, h, tol, dHat2, kappa, mu, epsv2, fric_iters=1, comp_range=None,
                         mu_comp=None, thickness=0.0, lin_rtol=1e-12, max_newton=0, dbc_idx=None, dbc_target=None,
                         body_force=None, flow=False):
        """One step with lagged friction (mu, epsv2, fricIterAmt, and optional
        per-component coefficients compNodeRange / muComp)."""
        X = _f(X).copy()
        vel = _f(vel).copy()
        bf = None if body_force is None else _f(body_force).copy()
        kap = np.ascontiguousarray(kappa, dtype=np.float64).copy()
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        tgt = None if dbc_target is None else _f(dbc_target)
        cr = None if comp_range is None else _i(comp_range)
        mc = None if mu_comp is None else _f(mu_comp)
        st = Stats()
        it = lib().orc_advance_friction(
            self.nV, X, self.X0, vel, self.mass, self.nT, self.tets, self.IB, self.vol, self.lam, self.mu,
            self.bnode.size, self.bnode, self.bedge.shape[0], _i(self.bedge), self.btri.shape[0], _i(self.btri),
            dbc_idx.size, dbc_idx, None if tgt is None else tgt.ctypes.data_as(C.c_void_p), h, tol, 1, dHat2, kap,
            thickness, lin_rtol, max_newton, None if bf is None else bf.ctypes.data_as(C.c_void_p), int(bool(flow)),
            mu, epsv2, int(fric_iters), 0 if cr is None else cr.size,
            None if cr is None else cr.ctypes.data_as(C.c_void_p), None if mc is None else mc.ctypes.data_as(C.c_void_p),
            C.byref(st))
        return it, X.reshape(-1, 3), vel.reshape(-1, 3), kap, st.as_dict()

    def advance(self, X, vel, h, tol, dHat2, kappa, thickness=0.0, lin_rtol=1e-12, max_newton=0,
                with_collision=True, dbc_idx=None, dbc_target=None, body_force=None, flow=False):
        """One step; dbc_target (n, 3) moves the DBC nodes dbc_idx (None: they stay);
        body_force (nV, 3) or None; flow: the Advance_One_Step_IE_Flow variant."""
        X = _f(X).copy()
        vel = _f(vel).copy()
        bf = None if body_force is None else _f(body_force).copy()
        kap = np.ascontiguousarray(kappa, dtype=np.float64).copy()
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        tgt = None if dbc_target is None else _f(dbc_target)
        st = Stats()
        it = lib().orc_advance_one_step(self.nV, X, self.X0, vel, self.mass, self.nT, self.tets, self.IB, self.vol,
                                        self.lam, self.mu, self.bnode.size, self.bnode, self.bedge.shape[0],
                                        _i(self.bedge), self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx,
                                        None if tgt is None else tgt.ctypes.data_as(C.c_void_p),
                                        h, tol, int(with_collision), dHat2, kap, thickness, lin_rtol, max_newton,
                                        None if bf is None else bf.ctypes.data_as(C.c_void_p), int(bool(flow)),
                                        C.byref(st))
        return it, X.reshape(-1, 3), vel.reshape(-1, 3), kap, st.as_dict()


class FrictionSet:
    """Compute_Friction_Basis at X on the barrier constraint set cs: fcs (n, 4),
    closest-point parameters cp (n, 2), tangent bases (n, 3, 2) and normal
    forces (n)."""

    def __init__(self, X, cs, dHat2, kappa, thickness=0.0):
        cs = _i(cs)
        n = cs.size // 4
        self.fcs = np.zeros(4 * n + 4, np.int32)
        self.cp = np.zeros(2 * n + 2)
        self.basis = np.zeros(6 * n + 6)
        self.nforce = np.zeros(n + 1)
        m = lib().orc_friction_basis(_f(X), n, cs, dHat2, kappa, thickness, self.fcs, self.cp, self.basis,
                                     self.nforce)
        self.n = int(m)
        self.fcs = self.fcs[: 4 * m].copy()
        self.cp = self.cp[: 2 * m].copy()
        self.basis = self.basis[: 6 * m].copy()
        self.nforce = self.nforce[:m].copy()

    def tangent_basis(self):
        """(n, 3, 2): columns of each constraint's tangent basis."""
        return self.basis.reshape(-1, 2, 3).transpose(0, 2, 1)

    def coef(self, comp_range, mu_comp):
        cr = _i(comp_range)
        rc = lib().orc_friction_coef(self.n, self.fcs, cr.size, cr, _f(mu_comp), self.nforce)
        if rc:
            raise ValueError("a friction vertex is in no component")

    def energy(self, X, Xn, epsvh2, mu):
        return lib().orc_friction_energy(_f(X), _f(Xn), self.n, self.fcs, self.cp, self.basis, self.nforce,
                                         epsvh2, mu)

    def gradient(self, X, Xn, epsvh2, mu):
        g = np.zeros(_f(X).size)
        lib().orc_friction_gradient(_f(X), _f(Xn), self.n, self.fcs, self.cp, self.basis, self.nforce, epsvh2, mu,
                                    g)
        return g

    def hessian_local(self, X, Xn, epsvh2, mu, project=True):
        H = np.zeros(144 * self.n + 1)
        vid = np.zeros(4 * self.n + 1, np.int32)
        lib().orc_friction_hessian_local(_f(X), _f(Xn), self.n, self.fcs, self.cp, self.basis, self.nforce, epsvh2,
                                         mu, int(project), H, vid)
        return H[: 144 * self.n].reshape(self.n, 144), vid[: 4 * self.n].reshape(self.n, 4)


def set_accd_max_steps(n):
    """ACCD step cap of every later CCD call of this process (<= 0: none, the default)."""
    lib().orc_set_accd_max_steps(int(n))


def geom_probe(kind, pts):
    """(type, unclassified squared distance) of a PT / EE / PE configuration."""
    x = np.zeros(12)
    pts = np.asarray(pts, dtype=np.float64).ravel()
    x[: pts.size] = pts
    t = C.c_int()
    d = C.c_double()
    lib().orc_geom_probe(int(kind), x, C.byref(t), C.byref(d))
    return t.value, d.value


def barrier_d2b(d, dh, k):
    t = d - dh
    return k * (-2.0 * np.log(d / dh) - 4.0 * t / d + t * t / (d * d))


def initialize_oipc(dhat, mass, stiff_mult=1.0):
    """Initialize_EIPC<double,false> (Library/FEM/Shell/DISCRETE_SHELL.h:555-577)."""
    dHat2 = dhat * dhat
    Hb = barrier_d2b(1.0e-16, dHat2, 1.0)
    k0 = stiff_mult * 1.0e11 * float(np.mean(mass)) * 3 / (4.0e-16 * Hb)
    return dHat2, np.array([k0, 100 * k0, 0.0])


def _vp(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dihedral(x4, grad=False):
    """(angle, jet gradient (12,) or None) of the hinge x4 = (x0, x1, x2, x3)."""
    x = _f(x4)
    g = np.zeros(12) if grad else None
    th = lib().orc_dihedral(x, _vp(g))
    return th, g


class ShellMesh:
    """A shell (triangles, no tets) as Initialize_Discrete_Shell sets it up,
    with the hinge stencils and rest data of Compute_Discrete_Shell_Inv_Basis;
    bending stiffness k = bend_mult E t^3 / (24 (1 - nu^2))."""

    def __init__(self, X, tris, density, thickness, E=0.0, nu=0.0, bend_mult=1.0):
        self.X = _f(X).copy()
        self.X0 = self.X.copy()
        tris = _i(tris)
        self.nV = self.X.size // 3
        nE = tris.size // 3
        kept = np.zeros(3 * nE, np.int32)
        self.mass = np.zeros(self.nV)
        self.tvol = np.zeros(nE)
        self.lam = np.zeros(nE)
        self.mu = np.zeros(nE)
        n = lib().orc_shell_init(self.nV, self.X, nE, tris, density, E, nu, thickness, kept, self.mass, self.tvol,
                                 self.lam, self.mu)
        self.nE = n
        self.tris = kept[: 3 * n].copy()
        self.tvol, self.lam, self.mu = self.tvol[:n].copy(), self.lam[:n].copy(), self.mu[:n].copy()
        bn = np.zeros(self.nV, np.int32)
        be = np.zeros(6 * n, np.int32)
        bt = np.zeros(3 * n, np.int32)
        nb, ne, nf = C.c_int(), C.c_int(), C.c_int()
        lib().orc_shell_surface(self.nV, self.X, 0, np.zeros(0, np.int32), n, self.tris, C.byref(nb), C.byref(ne),
                                C.byref(nf), bn, be, bt)
        self.bnode = bn[: nb.value].copy()
        self.bedge = be[: 2 * ne.value].reshape(-1, 2).copy()
        self.btri = bt[: 3 * nf.value].reshape(-1, 3).copy()
        self.dbc = np.zeros(self.nV, np.uint8)
        self.tIB = np.zeros(3 * n)
        lib().orc_shell_rest_fff(self.X0, n, self.tris, self.tIB)
        st = np.zeros(4 * max(self.bedge.shape[0], 1), np.int32)
        rest = np.zeros(3 * max(self.bedge.shape[0], 1))
        nh = lib().orc_shell_hinges(self.X0, n, self.tris, self.bedge.shape[0], _i(self.bedge), st, rest)
        self.nH = nh
        self.hinge = st[: 4 * nh].copy()
        self.hrest = rest[: 3 * nh].copy()
        self.bend_k = bend_mult * E * thickness ** 3 / (24 * (1 - nu * nu))

    ccd = Mesh.ccd  # the same surface step size (nodes, edges and triangles of the shell)

    def set_dbc(self, idx):
        self.dbc[:] = 0
        self.dbc[np.asarray(idx, dtype=np.int64)] = 1

    def elastic_terms(self, X, h, project=True, want=("E", "g", "H")):
        """Membrane + bending: (E, g (3 nV), Hm (nE, 9, 9), Hb (nH, 12, 12))."""
        E = C.c_double()
        g = np.zeros(3 * self.nV) if "g" in want else None
        Hm = np.zeros(81 * self.nE) if "H" in want else None
        Hb = np.zeros(144 * max(self.nH, 1)) if "H" in want else None
        lib().orc_shell_elastic_terms(self.nV, _f(X), self.nE, self.tris, self.tIB, self.tvol, self.lam, self.mu,
                                      self.nH, self.hinge, self.hrest, self.bend_k, h * h, _vp(self.dbc),
                                      int(project), C.cast(C.byref(E), C.c_void_p), _vp(g), _vp(Hm), _vp(Hb))
        return (E.value, g, None if Hm is None else Hm.reshape(self.nE, 9, 9),
                None if Hb is None else Hb[: 144 * self.nH].reshape(self.nH, 12, 12))

    def newton_system_hinge(self, X, Xtilde, h, dHat2, kappa, thickness=0.0, dbc_idx=None):
        """(rowptr, col, val, g) of the Hinge variant's projected system at X."""
        X = _f(X)
        Xt = _f(Xtilde)
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        cap = 64 * self.nV + 1024
        while True:
            rp = np.zeros(self.nV + 1, np.int32)
            col = np.zeros(cap, np.int32)
            val = np.zeros(9 * cap)
            g = np.zeros(3 * self.nV)
            n = lib().orc_newton_system_shell(self.nV, X, Xt, self.X0, self.mass, self.nE, self.tris, self.tvol,
                                              self.tIB, self.lam, self.mu, self.nH, self.hinge, self.hrest,
                                              self.bend_k, self.bnode.size, self.bnode, self.bedge.shape[0],
                                              _i(self.bedge), self.btri.shape[0], _i(self.btri), dbc_idx.size,
                                              dbc_idx, h, dHat2, kappa, thickness, rp, col, val, g, cap)
            if n >= 0:
                return rp, col[:n], val[: 9 * n].reshape(n, 3, 3), g
            cap = -n

    def advance_hinge(self, X, vel, h, tol, dHat2, kappa, thickness=0.0, lin_rtol=1e-12, max_newton=0,
                      with_collision=True, dbc_idx=None, body_force=None):
        """One Advance_One_Step_IE_Hinge step (DBC nodes stay)."""
        X = _f(X).copy()
        vel = _f(vel).copy()
        bf = None if body_force is None else _f(body_force).copy()
        kap = np.ascontiguousarray(kappa, dtype=np.float64).copy()
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        st = Stats()
        it = lib().orc_advance_shell(self.nV, X, self.X0, vel, self.mass, self.nE, self.tris, self.tvol, self.tIB,
                                     self.lam, self.mu, self.nH, self.hinge, self.hrest, self.bend_k,
                                     self.bnode.size, self.bnode, self.bedge.shape[0], _i(self.bedge),
                                     self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx, h, tol,
                                     int(with_collision), dHat2, kap, thickness, lin_rtol, max_newton,
                                     None if bf is None else bf.ctypes.data_as(C.c_void_p), C.byref(st))
        return it, X.reshape(-1, 3), vel.reshape(-1, 3), kap, st.as_dict()

    def advance_hinge_friction(self, X, vel, h, tol, dHat2, kappa, mu, epsv2, fric_iters=1, thickness=0.0,
                               lin_rtol=1e-12, max_newton=0, dbc_idx=None, body_force=None):
        """One Advance_One_Step_IE_Hinge step with lagged friction (DBC nodes stay)."""
        X = _f(X).copy()
        vel = _f(vel).copy()
        bf = None if body_force is None else _f(body_force).copy()
        kap = np.ascontiguousarray(kappa, dtype=np.float64).copy()
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        st = Stats()
        it = lib().orc_advance_shell_friction(
            self.nV, X, self.X0, vel, self.mass, self.nE, self.tris, self.tvol, self.tIB, self.lam, self.mu,
            self.nH, self.hinge, self.hrest, self.bend_k, self.bnode.size, self.bnode, self.bedge.shape[0],
            _i(self.bedge), self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx, h, tol, dHat2, kap, thickness,
            lin_rtol, max_newton, None if bf is None else bf.ctypes.data_as(C.c_void_p), mu, epsv2, int(fric_iters),
            C.byref(st))
        return it, X.reshape(-1, 3), vel.reshape(-1, 3), kap, st.as_dict()

    def normal_flow_force(self, X, magnitude):
        b = np.zeros(3 * self.nV)
        lib().orc_normal_flow_force(self.nV, _f(X), self.nE, self.tris, self.mass, magnitude, b)
        return b.reshape(-1, 3)

    def flow_terms(self, X, h):
        """(E, g, H (nE, 3, 3, 3, 3)) of the flow term alone."""
        E = C.c_double()
        g = np.zeros(3 * self.nV)
        H = np.zeros(81 * self.nE)
        lib().orc_flow_terms(self.nV, _f(X), self.nE, self.tris, self.tvol, h, C.cast(C.byref(E), C.c_void_p),
                             g.ctypes.data_as(C.c_void_p), H.ctypes.data_as(C.c_void_p))
        return E.value, g, H.reshape(self.nE, 3, 3, 3, 3)

    def constraint_set(self, X, dHat2, thickness=0.0):
        return Mesh.constraint_set(self, X, dHat2, thickness)

    def newton_system(self, X, Xtilde, h, dHat2, kappa, thickness=0.0, dbc_idx=None):
        """(rowptr, col, val (nnzb,3,3), g) of the flow variant's projected system at X."""
        X = _f(X)
        Xt = _f(Xtilde)
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        cap = 64 * self.nV + 1024
        while True:
            rp = np.zeros(self.nV + 1, np.int32)
            col = np.zeros(cap, np.int32)
            val = np.zeros(9 * cap)
            g = np.zeros(3 * self.nV)
            n = lib().orc_newton_system_shell_flow(self.nV, X, Xt, self.X0, self.mass, self.nE, self.tris, self.tvol,
                                                   self.bnode.size, self.bnode, self.bedge.shape[0], _i(self.bedge),
                                                   self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx, h,
                                                   dHat2, kappa, thickness, rp, col, val, g, cap)
            if n >= 0:
                return rp, col[:n], val[: 9 * n].reshape(n, 3, 3), g
            cap = -n

    def advance(self, X, vel, h, tol, dHat2, kappa, thickness=0.0, lin_rtol=1e-12, max_newton=0,
                with_collision=True, dbc_idx=None, body_force=None):
        """One Advance_One_Step_IE_Flow step (DBC nodes stay)."""
        X = _f(X).copy()
        vel = _f(vel).copy()
        bf = None if body_force is None else _f(body_force).copy()
        kap = np.ascontiguousarray(kappa, dtype=np.float64).copy()
        dbc_idx = np.zeros(0, np.int32) if dbc_idx is None else _i(dbc_idx)
        st = Stats()
        it = lib().orc_advance_shell_flow(self.nV, X, self.X0, vel, self.mass, self.nE, self.tris, self.tvol,
                                          self.bnode.size, self.bnode, self.bedge.shape[0], _i(self.bedge),
                                          self.btri.shape[0], _i(self.btri), dbc_idx.size, dbc_idx, h, tol,
                                          int(with_collision), dHat2, kap, thickness, lin_rtol, max_newton,
                                          None if bf is None else bf.ctypes.data_as(C.c_void_p), C.byref(st))
        return it, X.reshape(-1, 3), vel.reshape(-1, 3), kap, st.as_dict()
