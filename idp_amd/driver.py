"""Tet animation-fix driver on top of the C ABI.

Mirrors the calls the reference's Python driver makes for a tet scene
(``Python/Drivers/FEMDiscreteShellBase.py:214-221, 239-243, 297-361``):
``initialize_added_objects`` (volumes, inverse bases, Lame parameters, lumped
mass), ``initialize_OIPC`` (dhat^2 and kappa), the target expressed as the
velocity ``(x_target - x)/h`` (``MeshIO.Load_Velocity_X0``) and
``Advance_One_Step_IE_Hinge`` per step.
"""
from __future__ import annotations

import numpy as np

from . import capi


class TetAnimFix:
    def __init__(self, scene, device=0, pcg_rtol=1e-8, kappa_mult=1.0):
        self.scene = scene
        X = np.ascontiguousarray(scene.X, dtype=np.float64)
        tets = np.ascontiguousarray(scene.tets, dtype=np.int32)
        self.vol, self.IB, self.lam, self.mu, self.mass = capi.fem_init(X, tets, scene.density, scene.E, scene.nu)
        self.ctx = capi.Context(device)
        self.ctx.set_mesh(X, tets, self.IB, self.vol, self.lam, self.mu, self.mass)
        self.dHat2, self.kappa = capi.initialize_oipc(scene.dhat, self.mass, kappa_mult)
        self.h = scene.h
        self.tol = scene.newton_tol
        self.pcg_rtol = pcg_rtol
        self.ctx.set_state(X, scene.velocity)
        self.pn_iters = 0

    def set_target(self, target):
        """Load_Velocity_X0: v = (target - x) / h, formed on the device."""
        self.ctx.load_velocity(target, self.h)

    def advance_one_time_step(self, max_newton=0, verbose=0, body_force=None, flow=False):
        """Advance_One_Step_IE_Hinge (flow=False) or Advance_One_Step_IE_Flow (flow=True)."""
        it, st = self.ctx.advance(self.h, self.tol, self.dHat2, self.kappa, 0.0, True, self.pcg_rtol,
                                  max_newton=max_newton, verbose=verbose, body_force=body_force, flow=flow)
        self.pn_iters += it
        return it, st

    def state(self):
        return self.ctx.get_state()

    def close(self):
        self.ctx.close()


class ShellNormalFlow:
    """Normal-flow shell driver: Initialize_Shell_Hinge_EIPC's set-up
    (FEMDiscreteShellBase.py:190-203), initialize_OIPC, and per step the
    normal-flow force (Update_Normal_Flow_Neumann) followed by
    Advance_One_Step_IE_Flow (FEMDiscreteShellBase.py:268-273, 332-341).  The
    force is formed on the device from the step's start positions."""

    def __init__(self, scene, device=0, pcg_rtol=1e-8, kappa_mult=1.0):
        self.scene = scene
        X = np.ascontiguousarray(scene.X, dtype=np.float64)
        self.tris, self.mass, self.tvol, _, _ = capi.shell_init(X, scene.tris, scene.density, 0.0, 0.0,
                                                                  scene.thickness)
        self.ctx = capi.Context(device)
        self.ctx.set_mesh_shell(X, self.mass, self.tris, self.tvol)
        self.ctx.set_normal_flow(scene.magnitude)
        self.dHat2, self.kappa = capi.initialize_oipc(scene.dhat, self.mass, kappa_mult)
        self.h = scene.h
        self.tol = scene.newton_tol
        self.pcg_rtol = pcg_rtol
        self.ctx.set_state(X, np.zeros_like(X))
        self.pn_iters = 0

    def advance_one_time_step(self, max_newton=0, verbose=0):
        it, st = self.ctx.advance(self.h, self.tol, self.dHat2, self.kappa, 0.0, True, self.pcg_rtol,
                                  max_newton=max_newton, verbose=verbose, flow=True)
        self.pn_iters += it
        return it, st

    def state(self):
        return self.ctx.get_state()

    def close(self):
        self.ctx.close()


class ShellHinge:
    """Shell driver of the non-flow step (Advance_One_Step_IE_Hinge): the set-up
    of Initialize_Shell_Hinge_EIPC (lumped mass, vol, Lame parameters, rest
    first fundamental forms, hinge stencils and the hinge stiffness
    E t^3 / (24 (1 - nu^2)) times bend_mult; FEMDiscreteShellBase.py:190-203),
    gravity as the body force m g, Dirichlet nodes that stay."""

    def __init__(self, X, tris, density, E, nu, thickness, h, newton_tol, dhat, gravity=(0.0, -9.81, 0.0),
                 bend_mult=1.0, dbc=None, device=0, pcg_rtol=1e-8, kappa_mult=1.0):
        X = np.ascontiguousarray(X, dtype=np.float64)
        self.tris, self.mass, self.tvol, self.lam, self.mu = capi.shell_init(X, tris, density, E, nu, thickness)
        self.hinges, self.hrest, self.tIB = capi.shell_hinges(X, self.tris)
        self.bend_k = bend_mult * E * thickness ** 3 / (24 * (1 - nu * nu))
        self.ctx = capi.Context(device)
        self.ctx.set_mesh_shell(X, self.mass, self.tris, self.tvol, hinges=self.hinges)
        self.ctx.set_shell_elasticity(self.tIB, self.lam, self.mu, self.hrest, self.bend_k)
        if dbc is not None and len(dbc):
            self.ctx.set_dbc(np.asarray(dbc, dtype=np.int32))
        self.body_force = self.mass[:, None] * np.asarray(gravity, dtype=np.float64)[None, :]
        self.dHat2, self.kappa = capi.initialize_oipc(dhat, self.mass, kappa_mult)
        self.h = h
        self.tol = newton_tol
        self.pcg_rtol = pcg_rtol
        self.ctx.set_state(X, np.zeros_like(X))
        self.pn_iters = 0

    def advance_one_time_step(self, max_newton=0, verbose=0):
        it, st = self.ctx.advance(self.h, self.tol, self.dHat2, self.kappa, 0.0, True, self.pcg_rtol,
                                  max_newton=max_newton, verbose=verbose, body_force=self.body_force)
        self.pn_iters += it
        return it, st

    def state(self):
        return self.ctx.get_state()

    def close(self):
        self.ctx.close()


class BatchAnimFix:
    """Several independent animation-fix scenes advanced together in one
    context: every stage runs over all instances at once, while each instance
    keeps its own Newton, line-search and kappa sequence."""

    def __init__(self, scenes_, device=0, pcg_rtol=1e-8, kappa_mult=1.0):
        self.scenes = list(scenes_)
        meshes, dh2, kap, vel = [], [], [], []
        for sc in self.scenes:
            X = np.ascontiguousarray(sc.X, dtype=np.float64)
            tets = np.ascontiguousarray(sc.tets, dtype=np.int32)
            vol, IB, lam, mu, mass = capi.fem_init(X, tets, sc.density, sc.E, sc.nu)
            meshes.append((X, tets, IB, vol, lam, mu, mass))
            d2, k = capi.initialize_oipc(sc.dhat, mass, kappa_mult)
            dh2.append(d2)
            kap.append(k)
            vel.append(sc.velocity)
        self.ctx = capi.Context(device)
        self.ctx.set_mesh_batch(meshes)
        self.dHat2 = np.array(dh2)
        self.kappa = np.ascontiguousarray(np.stack(kap))
        self.h = self.scenes[0].h
        if any(sc.h != self.h for sc in self.scenes):
            raise ValueError("one time step for the whole batch")
        self.tol = np.array([sc.newton_tol for sc in self.scenes])
        self.pcg_rtol = pcg_rtol
        self.offsets = np.concatenate([[0], np.cumsum([sc.X.shape[0] for sc in self.scenes])])
        self.ctx.set_state(np.concatenate([sc.X for sc in self.scenes]), np.concatenate(vel))

    def advance_one_time_step(self, max_newton=0, verbose=0, accd_budget=None, pcg_handover=None):
        return self.ctx.advance_batch(self.h, self.tol, self.dHat2, self.kappa, 0.0, True, self.pcg_rtol,
                                      max_newton=max_newton, verbose=verbose, accd_budget=accd_budget,
                                      pcg_handover=pcg_handover)

    def state(self, i):
        X, v = self.ctx.get_state()
        a, b = self.offsets[i], self.offsets[i + 1]
        return X[a:b], v[a:b]

    def close(self):
        self.ctx.close()
