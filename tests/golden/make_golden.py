"""Generate golden vectors that pin the oracle independently of its code.

Every expected value comes from sympy: symbolic gradients and Hessians of the
closed-form squared distances (point-point, point-edge, point-triangle plane,
edge-edge line), the IPC log barrier b(d) = -k (d - dh)^2 ln(d/dh), the edge
mollifier e(x) = (2 - x/eps) x/eps, and the Neo-Hookean energy
psi(F) = mu/2 (|F|^2 - 3) - mu ln J + lam/2 ln^2 J, evaluated with mpmath at
40 digits.  PSD projections are the spectral clamp computed with
numpy.linalg.eigh (LAPACK); the clamp is unique, so the eigensolver used does
not matter.

Run:  python tests/golden/make_golden.py   (writes tests/golden/*.json)
"""
from __future__ import annotations

import json
import os

import mpmath
import numpy as np
import sympy as sp

mpmath.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def vec(name, n):
    return sp.Matrix(sp.symbols(f"{name}0:{n}", real=True))


def sqn(v):
    return (v.T * v)[0, 0]


def cross(a, b):
    return a.cross(b)


# ---------------------------------------------------------------- distances
X = sp.symbols("x0:12", real=True)
P0 = sp.Matrix(X[0:3])
P1 = sp.Matrix(X[3:6])
P2 = sp.Matrix(X[6:9])
P3 = sp.Matrix(X[9:12])


def d_pp(a, b):
    return sqn(a - b)


def d_pe(p, e0, e1):
    return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0)


def d_pt(p, t0, t1, t2):
    n = cross(t1 - t0, t2 - t0)
    return (((p - t0).T * n)[0, 0]) ** 2 / sqn(n)


def d_ee(a0, a1, b0, b1):
    n = cross(a1 - a0, b1 - b0)
    return (((b0 - a0).T * n)[0, 0]) ** 2 / sqn(n)


DH, KAP, MULT = sp.symbols("dh kappa mult", positive=True)


def barrier(d):
    return -KAP * (d - DH) ** 2 * sp.log(d / DH)


def moll(q, eps):
    r = q / eps
    return (2 - r) * r


def derivs(expr, nvars):
    vs = X[:nvars]
    g = [sp.diff(expr, v) for v in vs]
    H = [[sp.diff(gi, v) for v in vs] for gi in g]
    return g, H


def evaluate(expr, g, H, subs):
    f = sp.lambdify(list(subs.keys()), [expr, g, H], modules="mpmath")
    E, gv, Hv = f(*[mpmath.mpf(repr(float(v))) for v in subs.values()])
    return float(E), [float(x) for x in gv], [[float(x) for x in row] for row in Hv]


def psd(H):
    H = np.array(H)
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    if w.min() >= 0:
        return H.tolist()
    return ((V * np.maximum(w, 0)) @ V.T).tolist()


def barrier_cases(rng):
    """Constraint encodings of Library/FEM/IPC.h and their local Hessians."""
    eps_sym = sp.Symbol("eps", positive=True)
    kinds = {
        # name: (encoding builder over local vertices 0..3, nvars, energy expr)
        "PT": ((-1, 1, 2, 3), 12, barrier(d_pt(P0, P1, P2, P3))),
        "EE": ((0, 1, 2, 3), 12, barrier(d_ee(P0, P1, P2, P3))),
        "PE_x2": ((-1, 1, 2, -2), 9, MULT * barrier(d_pe(P0, P1, P2))),
        "PP_x3": ((-1, 1, -1, -3), 6, MULT * barrier(d_pp(P0, P1))),
        "mEE": ((0, 1, -3, 3), 12, barrier(d_ee(P0, P1, P2, P3)) * moll(sqn(cross(P1 - P0, P3 - P2)), eps_sym)),
        # mollified PE (p=ea0, edge eb0 eb1; other endpoint ea1): (ea0, eb0, eb1, -ea1-1)
        "mPE": ((0, 2, 3, -2), 12, barrier(d_pe(P0, P2, P3)) * moll(sqn(cross(P1 - P0, P3 - P2)), eps_sym)),
        # mollified PP (ea0, eb0, -ea1-1, -eb1-1)
        "mPP": ((0, 2, -2, -4), 12, barrier(d_pp(P0, P2)) * moll(sqn(cross(P1 - P0, P3 - P2)), eps_sym)),
    }
    cases = []
    for name, (enc, nv, expr) in kinds.items():
        g, H = derivs(expr, nv)
        for s in range(3):
            dh = 1e-2
            kappa = float(rng.uniform(1e3, 1e5))
            if name in ("PT",):
                t0, t1, t2 = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
                n = np.cross(t1 - t0, t2 - t0)
                n /= np.linalg.norm(n)
                w = rng.uniform(0.2, 0.3, 3)
                c = w[0] * t0 + w[1] * t1 + (1 - w[0] - w[1]) * t2
                p = c + n * rng.uniform(0.02, 0.08)
                Xv = np.concatenate([p, t0, t1, t2])
            elif name in ("EE",):
                a0, b0 = rng.normal(size=3), rng.normal(size=3)
                ua, ub = rng.normal(size=3), rng.normal(size=3)
                m = 0.5 * (a0 + b0)
                nrm = np.cross(ua, ub)
                nrm /= np.linalg.norm(nrm)
                a0 = m - 0.5 * ua
                a1 = m + 0.5 * ua
                b0 = m - 0.5 * ub + nrm * rng.uniform(0.02, 0.08)
                b1 = m + 0.5 * ub + nrm * rng.uniform(0.02, 0.08)
                Xv = np.concatenate([a0, a1, b0, b1])
            elif name == "PE_x2":
                e0, e1 = rng.normal(size=3), rng.normal(size=3)
                off = rng.normal(size=3)
                off -= off.dot(e1 - e0) / (e1 - e0).dot(e1 - e0) * (e1 - e0)
                off *= rng.uniform(0.02, 0.08) / np.linalg.norm(off)
                p = e0 + rng.uniform(0.3, 0.7) * (e1 - e0) + off
                Xv = np.concatenate([p, e0, e1, np.zeros(3)])
            elif name == "PP_x3":
                q = rng.normal(size=3)
                dvec = rng.normal(size=3)
                p = q + dvec / np.linalg.norm(dvec) * rng.uniform(0.02, 0.08)
                Xv = np.concatenate([p, q, np.zeros(6)])
            else:  # nearly parallel edge pairs for the mollified kinds
                a0 = rng.normal(size=3)
                u = rng.normal(size=3)
                u /= np.linalg.norm(u)
                perp = np.cross(u, rng.normal(size=3))
                perp /= np.linalg.norm(perp)
                a1 = a0 + u
                tilt = rng.normal(size=3) * 0.01
                b0 = a0 + perp * rng.uniform(0.03, 0.07) + u * rng.uniform(-0.2, 0.2)
                b1 = b0 + u + tilt
                Xv = np.concatenate([a0, a1, b0, b1])
            Xv = Xv.astype(float)
            X0 = Xv + rng.normal(size=12) * 0.05  # rest positions (mollifier threshold)
            a0r, a1r, b0r, b1r = X0[0:3], X0[3:6], X0[6:9], X0[9:12]
            eps = 1e-3 * np.sum((a0r - a1r) ** 2) * np.sum((b0r - b1r) ** 2)
            subs = {v: Xv[i] for i, v in enumerate(X[:nv])}
            subs.update({DH: dh, KAP: kappa})
            mult = {"PE_x2": 2.0, "PP_x3": 3.0}.get(name, 1.0)
            if name in ("PE_x2", "PP_x3"):
                subs[MULT] = mult
            if name.startswith("m"):
                subs[eps_sym] = eps
                q = np.sum(np.cross(Xv[3:6] - Xv[0:3], Xv[9:12] - Xv[6:9]) ** 2)
                assert q < eps, (name, q, eps)
            E, gv, Hv = evaluate(expr, g, H, subs)
            cases.append(dict(kind=name, cs=list(enc), X=Xv.tolist(), X0=X0.tolist(), dHat2=dh, kappa=kappa,
                              E=E, g=gv, H=Hv, Hproj=psd(Hv), nv=nv // 3))
    return cases


def nh_cases(rng):
    F = sp.Matrix(3, 3, sp.symbols("f0:9", real=True))  # row-major F(i,j) = f[3i+j]
    mu, lam = sp.symbols("mu lam", positive=True)
    J = F.det()
    psi = mu / 2 * (sum(F[i, j] ** 2 for i in range(3) for j in range(3)) - 3) - mu * sp.log(J) + lam / 2 * sp.log(J) ** 2
    fs = list(F)
    g = [sp.diff(psi, f) for f in fs]
    H = [[sp.diff(gi, f) for f in fs] for gi in g]
    fn = sp.lambdify([*fs, mu, lam], [psi, g, H], modules="mpmath")
    out = []
    for s in range(8):
        scale = [0.02, 0.1, 0.3, 0.6][s % 4]
        Fv = np.eye(3) + rng.normal(size=(3, 3)) * scale
        if np.linalg.det(Fv) <= 0.2:
            Fv = np.eye(3) + rng.normal(size=(3, 3)) * 0.05
        E, nu = float(rng.uniform(1e2, 1e5)), float(rng.uniform(0.1, 0.45))
        lv = E * nu / ((1 + nu) * (1 - 2 * nu))
        mv = E / (2 * (1 + nu))
        psiv, gv, Hv = fn(*[mpmath.mpf(repr(float(x))) for x in Fv.ravel()], mpmath.mpf(repr(mv)),
                          mpmath.mpf(repr(lv)))
        out.append(dict(F=Fv.ravel().tolist(), mu=mv, lam=lv, psi=float(psiv), P=[float(x) for x in gv],
                        H=[[float(x) for x in r] for r in Hv], Hproj=psd([[float(x) for x in r] for r in Hv])))
    return out


def barrier_scalar_cases():
    d, dh, k = sp.symbols("d dh k", positive=True)
    b = -k * (d - dh) ** 2 * sp.log(d / dh)
    fn = sp.lambdify([d, dh, k], [b, sp.diff(b, d), sp.diff(b, d, 2)], modules="mpmath")
    out = []
    for dv, dhv, kv in [(1e-4, 1e-3, 1.0), (5e-7, 1.08e-6, 4.9e5), (2.5e-3, 1e-2, 3e4), (1e-16, 1e-6, 1.0)]:
        vals = fn(mpmath.mpf(repr(dv)), mpmath.mpf(repr(dhv)), mpmath.mpf(repr(kv)))
        out.append(dict(d=dv, dh=dhv, k=kv, b=float(vals[0]), db=float(vals[1]), d2b=float(vals[2])))
    return out


def main():
    rng = np.random.default_rng(20240607)
    data = dict(barrier=barrier_cases(rng), neohookean=nh_cases(rng), barrier_scalar=barrier_scalar_cases())
    with open(os.path.join(HERE, "golden_derivatives.json"), "w") as f:
        json.dump(data, f)
    print({k: len(v) for k, v in data.items()})


if __name__ == "__main__":
    main()
