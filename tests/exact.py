"""Exact-arithmetic geometry used to pin the oracle (test helper).

Rational (fractions.Fraction) closest-feature distances for point-triangle,
point-segment and segment-segment pairs, and an exact orient3d-based
segment/triangle intersection test (floating-point filter with an exact
fallback) for the "every accepted iterate is intersection free" check.
"""
from __future__ import annotations

from fractions import Fraction as Fr

import numpy as np


def _f(v):
    return [Fr(float(c)) for c in v]


def _sub(a, b):
    return [a[i] - b[i] for i in range(3)]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def point_segment_d2(p, a, b):
    p, a, b = _f(p), _f(a), _f(b)
    ab, ap = _sub(b, a), _sub(p, a)
    if _dot(ab, ab) == 0:
        return _dot(ap, ap)
    t = _dot(ap, ab) / _dot(ab, ab)
    t = min(max(t, Fr(0)), Fr(1))
    q = [a[i] + t * ab[i] for i in range(3)]
    d = _sub(p, q)
    return _dot(d, d)


def point_triangle_d2(p, t0, t1, t2):
    """Exact squared distance (minimum over the face, edges and vertices)."""
    P, A, B, C = _f(p), _f(t0), _f(t1), _f(t2)
    n = _cross(_sub(B, A), _sub(C, A))
    # projection inside the triangle?
    s = _dot(_sub(P, A), n) / _dot(n, n)
    Q = [P[i] - s * n[i] for i in range(3)]
    inside = True
    for u, v in ((A, B), (B, C), (C, A)):
        if _dot(_cross(_sub(v, u), _sub(Q, u)), n) < 0:
            inside = False
    cands = [point_segment_d2(p, t0, t1), point_segment_d2(p, t1, t2), point_segment_d2(p, t2, t0)]
    if inside:
        d = _sub(P, Q)
        cands.append(_dot(d, d))
    return min(cands)


def segment_segment_d2(a0, a1, b0, b1):
    A0, A1, B0, B1 = _f(a0), _f(a1), _f(b0), _f(b1)
    u, v, w = _sub(A1, A0), _sub(B1, B0), _sub(A0, B0)
    a, b, c, d, e = _dot(u, u), _dot(u, v), _dot(v, v), _dot(u, w), _dot(v, w)
    D = a * c - b * b
    cands = [point_segment_d2(a0, b0, b1), point_segment_d2(a1, b0, b1), point_segment_d2(b0, a0, a1),
             point_segment_d2(b1, a0, a1)]
    if D != 0:
        s = (b * e - c * d) / D
        t = (a * e - b * d) / D
        if 0 <= s <= 1 and 0 <= t <= 1:
            pa = [A0[i] + s * u[i] for i in range(3)]
            pb = [B0[i] + t * v[i] for i in range(3)]
            dd = _sub(pa, pb)
            cands.append(_dot(dd, dd))
    return min(cands)


def _orient3d_exact(a, b, c, d):
    A, B, C, D = _f(a), _f(b), _f(c), _f(d)
    m = [_sub(A, D), _sub(B, D), _sub(C, D)]
    det = (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    return (det > 0) - (det < 0)


def orient3d(a, b, c, d):
    """Sign of det[a-d, b-d, c-d]: floating-point value accepted only when it
    exceeds Shewchuk's error bound (1e-14 x the permanent of the absolute
    values, looser than o3derrboundA), otherwise exact rational evaluation."""
    a, b, c, d = (np.asarray(x, dtype=np.float64) for x in (a, b, c, d))
    m = np.stack([a - d, b - d, c - d])
    m0, m1, m2 = m
    det = (m0[0] * (m1[1] * m2[2] - m1[2] * m2[1]) - m0[1] * (m1[0] * m2[2] - m1[2] * m2[0]) +
           m0[2] * (m1[0] * m2[1] - m1[1] * m2[0]))
    A = np.abs(m)
    perm = (A[0, 0] * (A[1, 1] * A[2, 2] + A[1, 2] * A[2, 1]) + A[0, 1] * (A[1, 0] * A[2, 2] + A[1, 2] * A[2, 0]) +
            A[0, 2] * (A[1, 0] * A[2, 1] + A[1, 1] * A[2, 0]))
    if abs(det) > 1e-14 * perm:
        return int(np.sign(det))
    return _orient3d_exact(a, b, c, d)


def segment_hits_triangle(e0, e1, t0, t1, t2):
    """True if the closed segment properly crosses the triangle (exact predicates)."""
    s0 = orient3d(t0, t1, t2, e0)
    s1 = orient3d(t0, t1, t2, e1)
    if s0 == s1 or s0 == 0 and s1 == 0:
        return False
    a = orient3d(e0, e1, t0, t1)
    b = orient3d(e0, e1, t1, t2)
    c = orient3d(e0, e1, t2, t0)
    return (a >= 0 and b >= 0 and c >= 0) or (a <= 0 and b <= 0 and c <= 0)


def surface_self_intersects(X, edges, tris):
    """Exact edge-triangle intersection test over AABB-overlapping, non-adjacent pairs."""
    X = np.asarray(X).reshape(-1, 3)
    E = np.asarray(edges)
    T = np.asarray(tris)
    elo = np.minimum(X[E[:, 0]], X[E[:, 1]])
    ehi = np.maximum(X[E[:, 0]], X[E[:, 1]])
    tlo = np.minimum(np.minimum(X[T[:, 0]], X[T[:, 1]]), X[T[:, 2]])
    thi = np.maximum(np.maximum(X[T[:, 0]], X[T[:, 1]]), X[T[:, 2]])
    for i in range(E.shape[0]):
        ov = np.all((elo[i] <= thi) & (tlo <= ehi[i]), axis=1)
        for f in np.nonzero(ov)[0]:
            t = T[f]
            if E[i, 0] in t or E[i, 1] in t:
                continue
            if segment_hits_triangle(X[E[i, 0]], X[E[i, 1]], X[t[0]], X[t[1]], X[t[2]]):
                return True
    return False
