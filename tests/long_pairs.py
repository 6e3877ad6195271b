"""Two-triangle scenes around the long additive-CCD pairs of
tests/golden/accd_long_pairs.json (start positions and motions of the pair's
four vertices, and the bound it started with).

An edge-edge pair becomes two triangles, each edge with a third vertex half
an edge length away on the far side of the plane the two edges nearly share
(so the other pairs of the scene are far apart); a point-triangle pair
becomes the triangle and a small triangle standing on the point, away from
the triangle's plane.  The third vertices move with the mean of their edge's
(or the point's) motion."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return json.load(open(os.path.join(HERE, "golden", "accd_long_pairs.json")))["pairs"]


def scene(pair):
    x = np.array(pair["x"]).reshape(4, 3)
    p = np.array(pair["p"]).reshape(4, 3)
    if pair["ee"]:
        a0, a1, b0, b1 = x
        u, v = a1 - a0, b1 - b0
        c = np.cross(u, v)
        c /= np.linalg.norm(c)
        if np.dot(c, b0 - a0) < 0:
            c = -c
        L = max(np.linalg.norm(u), np.linalg.norm(v))
        X = np.array([a0, a1, 0.5 * (a0 + a1) - 0.5 * L * c, b0, b1, 0.5 * (b0 + b1) + 0.5 * L * c])
        P = np.array([p[0], p[1], 0.5 * (p[0] + p[1]), p[2], p[3], 0.5 * (p[2] + p[3])])
    else:
        q, t0, t1, t2 = x
        n = np.cross(t1 - t0, t2 - t0)
        n /= np.linalg.norm(n)
        if np.dot(n, q - t0) < 0:
            n = -n
        e = (t1 - t0) / np.linalg.norm(t1 - t0)
        L = np.linalg.norm(t1 - t0)
        X = np.array([t0, t1, t2, q, q + 0.5 * L * n + 0.3 * L * e, q + 0.5 * L * n - 0.3 * L * e])
        P = np.array([p[1], p[2], p[3], p[0], p[0], p[0]])
    return X, np.array([[0, 1, 2], [3, 4, 5]], np.int32), P


def sliding_point(normal_speed, d=1e-8, thickness=0.0):
    """A point at distance thickness + d above a unit triangle, moving 1e-3
    along it and normal_speed * d along the normal over the step: no hit for
    normal_speed > -0.9, a hit near t = 0.9 / -normal_speed below."""
    t = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    q = np.array([0.25, 0.25, thickness + d])
    x = np.vstack([q, t])
    p = np.zeros((4, 3))
    p[0] = [1e-3, 0.5e-3, normal_speed * d]
    # turned off the axes, so that the boxes of the swept AABB test overlap
    R, _ = np.linalg.qr(np.array([[0.8, 0.3, -0.5], [0.1, 0.9, 0.4], [0.6, -0.2, 0.7]]))
    return {"ee": False, "alpha": 1.0, "x": (x @ R.T + 0.3).ravel().tolist(), "p": (p @ R.T).ravel().tolist()}
