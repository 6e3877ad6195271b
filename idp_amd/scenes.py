"""Synthetic tet scenes for the animation-fix workloads.

Every mesh is a block of unit cells, each split into the 6-tet Kuhn template of
``Projects/FEM/input/cube.vtk:5-21`` (all six tets share the 0-6 diagonal, so
translated copies are conforming).  Targets are analytic maps of the rest
positions; the solver receives them as the velocity ``v = (x_target - x)/h``,
the ``Load_Velocity_X0`` semantics of ``Library/Utils/MESHIO.h:987-1005``.

Material and solver constants follow the reference animation-fix driver
``Projects/FEMShell/16_fix_char_seq.py:27-37`` (dt 0.04, density 1000,
E 100, nu 0.4, zero gravity, kappa multiplier 1), with dhat and the Newton
tolerance relative to the rest bounding-box diagonal as the workloads state.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# cube.vtk corner offsets (point order of the file) and its six tets
_CORNERS = np.array(
    [[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1], [0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]], dtype=np.int64
)
_KUHN = np.array(
    [[0, 4, 6, 5], [3, 6, 2, 0], [0, 4, 7, 6], [3, 6, 0, 7], [2, 0, 6, 1], [6, 0, 5, 1]], dtype=np.int64
)

DT = 0.04
DENSITY = 1000.0
YOUNGS = 100.0
POISSON = 0.4


@dataclasses.dataclass
class Scene:
    name: str
    X: np.ndarray  # (V,3) float64 start positions (= rest)
    tets: np.ndarray  # (T,4) int32
    target: np.ndarray  # (V,3) float64
    h: float = DT
    density: float = DENSITY
    E: float = YOUNGS
    nu: float = POISSON
    dhat_rel: float = 1e-3
    tol_rel: float = 1e-4

    @property
    def diag(self) -> float:
        return float(np.linalg.norm(self.X.max(0) - self.X.min(0)))

    @property
    def dhat(self) -> float:
        return self.dhat_rel * self.diag

    @property
    def newton_tol(self) -> float:
        return self.tol_rel * self.diag

    @property
    def velocity(self) -> np.ndarray:
        return (self.target - self.X) / self.h


def kuhn_block(nx: int, ny: int, nz: int, spacing: float):
    """Regular (nx+1)(ny+1)(nz+1) grid, vertex id = i + (nx+1)(j + (ny+1)k)."""
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    X = np.stack([ii.ravel(order="F"), jj.ravel(order="F"), kk.ravel(order="F")], 1).astype(np.float64)
    X *= spacing
    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")], 1)
    corner = cells[:, None, :] + _CORNERS[None, :, :]  # (C,8,3)
    cid = corner[..., 0] + (nx + 1) * (corner[..., 1] + (ny + 1) * corner[..., 2])
    tets = cid[:, _KUHN].reshape(-1, 4)
    return X, tets.astype(np.int32)


def tet_volumes(X, tets):
    a = X[tets[:, 1]] - X[tets[:, 0]]
    b = X[tets[:, 2]] - X[tets[:, 0]]
    c = X[tets[:, 3]] - X[tets[:, 0]]
    return np.einsum("ij,ij->i", np.cross(a, b), c) / 6.0


def _jitter(X, sigma, seed):
    rng = np.random.default_rng(seed)
    return X + rng.normal(0.0, sigma, X.shape)


def _bend_with_legs(X, x_mid, y_mid, radius, angle_deg):
    """Bend about the z axis: the middle of the bar (arc length radius*angle)
    wraps onto a circle of the given centre-line radius, the rest continues as
    straight tangent legs, so a bend beyond ~180 degrees makes the legs cross."""
    s = X[:, 0] - x_mid
    t = X[:, 1] - y_mid
    half = math.radians(angle_deg) / 2
    a = radius * half
    out = X.copy()
    arc = np.abs(s) <= a
    phi = np.where(arc, s / radius, np.sign(s) * half)
    rr = radius - t
    bx = rr * np.sin(phi)
    by = radius - rr * np.cos(phi)
    leg = np.where(arc, 0.0, np.abs(s) - a)
    sg = np.sign(s)
    out[:, 0] = bx + leg * sg * np.cos(phi)
    out[:, 1] = by + leg * sg * np.sin(phi)
    return out


def bar_bend(nx=20, ny=4, nz=4, spacing=0.05, radius=0.15, angle_deg=220.0, jitter=0.0, seed=0,
             name="bar_bend_220"):
    """Bar of nx*ny*nz cells bent about a circle of the given radius."""
    X, tets = kuhn_block(nx, ny, nz, spacing)
    if jitter > 0:
        X = _jitter(X, jitter * spacing, seed)
    L, W = nx * spacing, ny * spacing
    target = _bend_with_legs(X, L / 2, W / 2, radius, angle_deg)
    return Scene(name, X, tets, target)


def bar_twist_fold(nx=64, ny=16, nz=16, spacing=0.05, twists=2.0, fold_radius=0.35, jitter=0.05, seed=1,
                   name="bar_twist2_fold180"):
    """Two full twists about the long axis, then a 180 degree fold at mid-length."""
    X, tets = kuhn_block(nx, ny, nz, spacing)
    X = _jitter(X, jitter * spacing, seed)
    L, W, D = nx * spacing, ny * spacing, nz * spacing
    ang = 2 * math.pi * twists * X[:, 0] / L
    y, z = X[:, 1] - W / 2, X[:, 2] - D / 2
    Y = np.stack([X[:, 0], W / 2 + np.cos(ang) * y - np.sin(ang) * z, D / 2 + np.sin(ang) * y + np.cos(ang) * z], 1)
    target = _bend_with_legs(Y, L / 2, W / 2, fold_radius, 180.0)
    return Scene(name, X, tets, target)


def sinusoid_target(X, side, n_modes, amp_range, wl_range, rng):
    disp = np.zeros_like(X)
    for _ in range(n_modes):
        k = rng.normal(size=3)
        k /= np.linalg.norm(k)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        A = rng.uniform(*amp_range) * side
        lam = rng.uniform(*wl_range) * side
        ph = rng.uniform(0, 2 * math.pi)
        disp += A * np.sin(2 * math.pi * (X @ k) / lam + ph)[:, None] * d[None, :]
    return X + disp


def cube_sinusoid(n, spacing=0.05, jitter=0.03, seed=0, n_modes=4, amp=(0.3, 0.6), wl=(0.5, 1.0)):
    X, tets = kuhn_block(n, n, n, spacing)
    rng = np.random.default_rng(seed)
    X = X + rng.normal(0.0, jitter * spacing, X.shape)
    target = sinusoid_target(X, n * spacing, n_modes, amp, wl, rng)
    return Scene(f"cube{n}_sin{n_modes}", X, tets, target)


def config1():
    """20x4x4 cells, spacing 0.05, 220 degree bend about radius 0.15 (ends overlap)."""
    return bar_bend()


def config2():
    """64x16x16 cells, jitter 0.05*spacing, two twists + 180 degree fold."""
    return bar_twist_fold()


def config3(count=256, seed=3):
    """count cubes of n^3 cells (n uniform in 4..16, seeded), jitter 0.03*spacing,
    targets from 4 seeded sinusoids (amplitude 0.3-0.6, wavelength 0.5-1.0 of the side)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(4, 17, size=count)
    return [cube_sinusoid(int(n), seed=int(seed * 100003 + i)) for i, n in enumerate(ns)]


def config5():
    """160x48x48 cells, jitter 0.02*spacing, crumple of 12 sinusoidal modes."""
    X, tets = kuhn_block(160, 48, 48, 0.05)
    rng = np.random.default_rng(5)
    X = X + rng.normal(0.0, 0.02 * 0.05, X.shape)
    width = 48 * 0.05
    target = sinusoid_target(X, width, 12, (0.1, 0.4), (0.5, 1.0), rng)
    return Scene("bar160x48x48_crumple12", X, tets, target)


def small_bend(nx=8, ny=2, nz=2, spacing=0.05, angle_deg=250.0, radius=0.06):
    """Tiny bar used by the smoke check and fast parity tests."""
    return bar_bend(nx, ny, nz, spacing, radius, angle_deg, name=f"bar{nx}x{ny}x{nz}_bend{int(angle_deg)}")

def two_blocks_push(n=3, spacing=0.05, gap=0.2, push=0.5):
    """Two n^3 Kuhn blocks side by side along x, gap (in cells) apart.  Moving
    Dirichlet nodes: block A's face toward block B is pushed toward B by
    `push` cells (past the gap when push > gap, so the CCD stops the move and
    the penalty takes over); block B's far face is held.  Returns (scene, dbc indices, dbc
    targets (n, 3)).  Every other vertex starts at rest with target = rest."""
    Xa, Ta = kuhn_block(n, n, n, spacing)
    Xb = Xa + np.array([(n + gap) * spacing, 0.0, 0.0])
    X = np.vstack([Xa, Xb])
    T = np.vstack([Ta, Ta + Xa.shape[0]]).astype(np.int32)
    face_a = np.nonzero(np.isclose(X[:, 0], n * spacing))[0]  # A's face toward B
    face_b = np.nonzero(np.isclose(X[:, 0], X[:, 0].max()))[0]
    idx = np.concatenate([face_a, face_b]).astype(np.int32)
    tgt = X[idx].copy()
    tgt[: face_a.size, 0] += push * spacing
    sc = Scene(f"two_blocks{n}_push", X.copy(), T, X.copy())
    return sc, idx, tgt


def two_blocks_slide(n=2, spacing=0.05, gap=2e-4, slide=1.0):
    """Two n^3 Kuhn blocks stacked along y, `gap` cells apart; the search
    direction moves both blocks down by 0.1 cell (so the swept boxes of the
    facing surfaces overlap) and slides the upper block by `slide` cells
    along x.  The facing surfaces stay `gap` apart for the whole motion, so
    the additive CCD walks each facing pair in steps of about 0.9 gap / slide
    of the step size.  Returns (scene, direction (V, 3))."""
    Xa, Ta = kuhn_block(n, n, n, spacing)
    Xb = Xa + np.array([0.0, (n + gap) * spacing, 0.0])
    X = np.vstack([Xa, Xb])
    T = np.vstack([Ta, Ta + Xa.shape[0]]).astype(np.int32)
    p = np.zeros_like(X)
    p[:, 1] = -0.1 * spacing
    p[Xa.shape[0]:, 0] = slide * spacing
    return Scene(f"two_blocks{n}_slide", X.copy(), T, X.copy()), p


def block_on_slab(n=2, spacing=0.05, gap_dhat=0.5, E=1e5, h=0.01, offset=0.5):
    """An n^3 Kuhn block resting on a (4n, 1, 2n)-cell slab, `gap_dhat` dhat
    above it, for friction: the slab's vertices are the Dirichlet nodes
    (returned as the third value), the block starts at rest, and a caller
    tilts gravity to make it slide or stick.  offset (cells, in x and z)
    shifts the block off the slab's grid so that its bottom vertices sit
    over slab triangles (point-triangle and edge-edge contacts) rather than
    over slab vertices (point-point contacts).  Returns (scene, slab vertex
    count, Dirichlet ids)."""
    Xs, Ts = kuhn_block(4 * n, 1, 2 * n, spacing)
    Xb, Tb = kuhn_block(n, n, n, spacing)
    sc0 = Scene("probe", np.vstack([Xs, Xb]), Ts, np.vstack([Xs, Xb]))
    dhat = sc0.dhat_rel * float(np.linalg.norm(np.array([4 * n, 1 + n, 2 * n]) * spacing))
    Xb = Xb + np.array([(1.5 * n + offset) * spacing, spacing + gap_dhat * dhat, (0.5 * n + offset) * spacing])
    X = np.vstack([Xs, Xb])
    T = np.vstack([Ts, Tb + Xs.shape[0]]).astype(np.int32)
    sc = Scene(f"block{n}_on_slab", X.copy(), T, X.copy(), h=h, E=E)
    return sc, Xs.shape[0], np.arange(Xs.shape[0], dtype=np.int32)


# ---------------------------------------------------------------------------
# Shell scenes for the normal-flow variant (Projects/FEMShell/12-14_normal_flow.py:
# dt 1 scaled by the smoothing intensity 0.5, normal-flow magnitude 0.005 /
# 0.5^2, density 1, thickness 1, E = nu = 0, zero gravity, dhat 1e-3,
# Newton tolerance PNTol 1e-3 of FEMDiscreteShellBase.py:43)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class ShellScene:
    name: str
    X: np.ndarray  # (V,3) float64 start positions
    tris: np.ndarray  # (F,3) int32, outward oriented
    h: float = 0.5
    magnitude: float = 0.02
    density: float = 1.0
    thickness: float = 1.0
    dhat: float = 1e-3
    newton_tol: float = 1e-3

    @property
    def diag(self) -> float:
        return float(np.linalg.norm(self.X.max(0) - self.X.min(0)))


def icosphere(level: int, radius: float = 1.0, center=(0.0, 0.0, 0.0)):
    """Icosahedron subdivided `level` times (each triangle into four at its
    edge midpoints), vertices pushed onto the sphere; triangles oriented so
    cross(X2 - X1, X3 - X1) points outward.  10 * 4^level + 2 vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[k] = len(V) - 1
            return mid[k]

        F2 = []
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            F2 += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = F2
    X = np.array(V) * radius + np.asarray(center, dtype=np.float64)
    tris = np.array(F, dtype=np.int32)
    # outward orientation (the icosahedron list above is outward already)
    c = np.cross(X[tris[:, 1]] - X[tris[:, 0]], X[tris[:, 2]] - X[tris[:, 0]])
    assert np.all(np.einsum("ij,ij->i", c, X[tris].mean(1) - np.asarray(center)) > 0)
    return X, tris


def sphere_pair(level=2, radius=0.5, gap=0.01, magnitude=0.2, name=None):
    """Two icospheres whose closest points are `gap` apart along x, with a
    normal-flow force strong enough to outgrow the flow's shrinking: both
    inflate until they press on each other."""
    return sphere_lattice((2, 1, 1), level, radius, gap, magnitude, name=name or f"spheres{level}_gap{gap:g}")


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def sphere_lattice(counts=(4, 4, 4), level=3, radius=0.5, gap=0.01, magnitude=0.2, seed=11, name=None):
    """counts[0] x counts[1] x counts[2] icospheres on a grid of pitch
    2 radius + gap, each turned by its own seeded rotation (so no vertex of
    one sits exactly opposite a vertex or edge of its neighbour); under the
    normal-flow force each presses on its neighbours (every sphere a separate
    shell component)."""
    rng = np.random.default_rng(seed)
    Xs, Ts, off = [], [], 0
    for k in range(counts[2]):
        for j in range(counts[1]):
            for i in range(counts[0]):
                c = np.array([i, j, k], dtype=np.float64) * (2 * radius + gap)
                X, T = icosphere(level, radius)
                X = X @ _rotation(rng).T + c
                Xs.append(X)
                Ts.append(T + off)
                off += X.shape[0]
    nm = name or f"spheres{counts[0]}x{counts[1]}x{counts[2]}_l{level}"
    return ShellScene(nm, np.vstack(Xs), np.vstack(Ts).astype(np.int32), magnitude=magnitude)


def shell_sheet(n=8, size=1.0, seed=0, amp=0.0):
    """n x n squares of side size/n in the x-z plane at y = 0, each cut into
    two triangles with normals along +y, optionally jittered along y."""
    xs = np.linspace(0.0, size, n + 1)
    X = np.array([[x, 0.0, z] for z in xs for x in xs])
    if amp > 0:
        X[:, 1] += np.random.default_rng(seed).normal(0, amp, X.shape[0])
    T = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            b, c, d = a + 1, a + n + 1, a + n + 2
            T += [[a, c, b], [b, c, d]]
    return X, np.array(T, np.int32)


def draped_sheet(n=8, level=2, radius=0.3, gap=0.003):
    """A sheet of n x n squares falling on a sphere shell (its nodes are
    Dirichlet) whose top is `gap` below the sheet's centre.  Returns
    (X, tris, dbc indices)."""
    Xc, Tc = shell_sheet(n)
    Xs, Ts = icosphere(level, radius, (0.5, -(radius + gap), 0.5))
    X = np.vstack([Xc, Xs])
    T = np.vstack([Tc, Ts + Xc.shape[0]]).astype(np.int32)
    return X, T, np.arange(Xc.shape[0], X.shape[0], dtype=np.int32)
