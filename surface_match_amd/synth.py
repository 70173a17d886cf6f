"""Seeded synthetic point clouds for the PPF detection workloads.

Model: a box [0,100]x[0,60]x[0,40] mm with a cylinder (r=15, axis +z at
(50,30), z in [40,90], top cap) and, optionally, a torus (R=30, r=8) lying
on the cylinder cap.  Points are area-uniform with analytic outward unit
normals.  Scenes place rigid copies under uniform SO(3) rotations
(Shoemake's quaternion method), cut each copy with a random half-space,
add a ground plane and/or uniform clutter, and iid Gaussian noise.

All randomness comes from ``numpy.random.Generator(PCG64(seed))`` so every
workload is reproducible bit for bit on any host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

VIEWPOINT = (620.0, 100.0, 500.0)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


# --------------------------------------------------------------------------
# model surface
# --------------------------------------------------------------------------

def _surfaces(torus: bool):
    """(name, area) list of the model's surface patches."""
    s = [
        ("z0", 100 * 60), ("z40", 100 * 60 - math.pi * 15 ** 2),
        ("y0", 100 * 40), ("y60", 100 * 40), ("x0", 60 * 40), ("x100", 60 * 40),
        ("cyl", 2 * math.pi * 15 * 50), ("cap", math.pi * 15 ** 2),
    ]
    if torus:
        s.append(("torus", 4 * math.pi ** 2 * 30 * 8))
    return s


def _sample_patch(name, m, rng):
    p = np.empty((m, 3))
    n = np.zeros((m, 3))
    u = rng.random(m)
    v = rng.random(m)
    if name == "z0":
        p[:] = np.stack([u * 100, v * 60, np.zeros(m)], 1); n[:, 2] = -1
    elif name == "z40":
        # top face minus the disk under the cylinder (rejection)
        out = []
        while sum(len(o) for o in out) < m:
            uu = rng.random(2 * m) * 100
            vv = rng.random(2 * m) * 60
            keep = (uu - 50) ** 2 + (vv - 30) ** 2 > 15 ** 2
            out.append(np.stack([uu[keep], vv[keep]], 1))
        q = np.concatenate(out)[:m]
        p[:] = np.stack([q[:, 0], q[:, 1], np.full(m, 40.0)], 1); n[:, 2] = 1
    elif name == "y0":
        p[:] = np.stack([u * 100, np.zeros(m), v * 40], 1); n[:, 1] = -1
    elif name == "y60":
        p[:] = np.stack([u * 100, np.full(m, 60.0), v * 40], 1); n[:, 1] = 1
    elif name == "x0":
        p[:] = np.stack([np.zeros(m), u * 60, v * 40], 1); n[:, 0] = -1
    elif name == "x100":
        p[:] = np.stack([np.full(m, 100.0), u * 60, v * 40], 1); n[:, 0] = 1
    elif name == "cyl":
        th = u * 2 * math.pi
        p[:] = np.stack([50 + 15 * np.cos(th), 30 + 15 * np.sin(th), 40 + 50 * v], 1)
        n[:] = np.stack([np.cos(th), np.sin(th), np.zeros(m)], 1)
    elif name == "cap":
        r = 15 * np.sqrt(u)
        th = v * 2 * math.pi
        p[:] = np.stack([50 + r * np.cos(th), 30 + r * np.sin(th), np.full(m, 90.0)], 1)
        n[:, 2] = 1
    elif name == "torus":
        R, r = 30.0, 8.0
        out = []
        while sum(len(o) for o in out) < m:
            a = rng.random(2 * m) * 2 * math.pi
            b = rng.random(2 * m) * 2 * math.pi
            w = rng.random(2 * m) * (R + r)
            keep = w < R + r * np.cos(b)
            out.append(np.stack([a[keep], b[keep]], 1))
        q = np.concatenate(out)[:m]
        a, b = q[:, 0], q[:, 1]
        c = np.array([50.0, 30.0, 90.0 + r])
        p[:] = c + np.stack([(R + r * np.cos(b)) * np.cos(a), (R + r * np.cos(b)) * np.sin(a),
                             r * np.sin(b)], 1)
        n[:] = np.stack([np.cos(b) * np.cos(a), np.cos(b) * np.sin(a), np.sin(b)], 1)
    return p, n


def make_model(n_points: int, seed: int, torus: bool = False):
    """Area-uniform model samples: (points f32 [n,3], normals f32 [n,3])."""
    rng = _rng(seed)
    surf = _surfaces(torus)
    area = np.array([a for _, a in surf])
    counts = rng.multinomial(n_points, area / area.sum())
    ps, ns = [], []
    for (name, _), m in zip(surf, counts):
        if m:
            p, n = _sample_patch(name, int(m), rng)
            ps.append(p)
            ns.append(n)
    p = np.concatenate(ps)
    n = np.concatenate(ns)
    perm = rng.permutation(len(p))
    return p[perm].astype(np.float32), n[perm].astype(np.float32)


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------

def random_rotation(rng) -> np.ndarray:
    """Uniform SO(3) rotation (Shoemake, Graphics Gems III, 1992)."""
    u1, u2, u3 = rng.random(3)
    a, b = math.sqrt(1 - u1), math.sqrt(u1)
    w, x, y, z = (b * math.cos(2 * math.pi * u3), a * math.sin(2 * math.pi * u2),
                  a * math.cos(2 * math.pi * u2), b * math.sin(2 * math.pi * u3))
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _cut(p, frac, rng):
    """Remove the fraction `frac` of points lowest along a random direction."""
    if frac <= 0:
        return p
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    s = (p - p.mean(0)) @ d
    thr = np.quantile(s, frac)
    return p[s >= thr]


@dataclass
class Scene:
    points: np.ndarray                 # f32 [N,3]
    gt_poses: list = field(default_factory=list)   # 4x4 float64, model -> scene
    view_point: tuple = VIEWPOINT
    info: dict = field(default_factory=dict)


def make_scene(seed: int, n_instances: int, pts_per_instance: int, occlusion: float,
               sigma: float, cube: float = 200.0, plane_points: int = 0,
               plane_z: float | None = None, clutter_frac: float = 0.0,
               torus: bool = False) -> Scene:
    rng = _rng(seed)
    parts, poses = [], []
    for k in range(n_instances):
        mp, _ = make_model(pts_per_instance, seed * 1000 + k + 1, torus)
        R = random_rotation(rng)
        t = rng.uniform(-cube, cube, 3)
        q = mp.astype(np.float64) @ R.T + t
        q = _cut(q, occlusion, rng)
        parts.append(q)
        T = np.eye(4)
        T[:3, :3] = R
        T[:3, 3] = t
        poses.append(T)
    inst = np.concatenate(parts) if parts else np.zeros((0, 3))
    extra = []
    if plane_points:
        z = -cube if plane_z is None else plane_z
        xy = rng.uniform(-cube, cube, (plane_points, 2))
        extra.append(np.concatenate([xy, np.full((plane_points, 1), z)], 1))
    if clutter_frac > 0 and len(inst):
        m = int(round(clutter_frac * len(inst) / (1 - clutter_frac)))
        lo, hi = inst.min(0), inst.max(0)
        extra.append(rng.uniform(lo, hi, (m, 3)))
    pts = np.concatenate([inst] + extra) if extra else inst
    pts = pts + rng.normal(scale=sigma, size=pts.shape)
    perm = rng.permutation(len(pts))
    return Scene(points=pts[perm].astype(np.float32), gt_poses=poses,
                 info=dict(instances=n_instances, n=len(pts)))


# --------------------------------------------------------------------------
# the benchmark workloads
# --------------------------------------------------------------------------

def workload(idx: int, scale: float = 1.0):
    """Model + scene for workload idx (1..5).  `scale` shrinks point counts
    for quick tests (1.0 = the full workload)."""
    base = 0x5EED0000 + idx
    if idx == 1:
        mp, mn = make_model(int(20000 * scale), base)
        sc = make_scene(base + 1, 3, int(20000 * scale), 0.3, 0.1, 200.0,
                        plane_points=int(50000 * scale))
        return dict(model=(mp, mn), scene=sc, train_rel=0.04, match=(0.04, 0.1, 0.2))
    if idx == 2:
        mp, mn = make_model(int(100000 * scale), base)
        sc = make_scene(base + 1, 10, int(133334 * scale), 0.4, 0.2, 300.0, clutter_frac=0.2)
        return dict(model=(mp, mn), scene=sc, train_rel=0.04, match=(0.04, 0.1, 0.2))
    if idx == 5:
        mp, mn = make_model(int(200000 * scale), base, torus=True)
        sc = make_scene(base + 1, 20, int(285715 * scale), 0.5, 0.3, 400.0,
                        clutter_frac=0.3, torus=True)
        return dict(model=(mp, mn), scene=sc, train_rel=0.04, match=(0.04, 0.1, 0.2))
    raise ValueError(idx)


def batch_scenes(seed: int, n_scenes: int = 64, lo: int = 50000, hi: int = 500000):
    """Scenes of uneven size (log-uniform point counts, 1..8 instances)."""
    rng = _rng(seed)
    out = []
    for s in range(n_scenes):
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        k = int(rng.integers(1, 9))
        plane = n // 4
        per = max(1000, (n - plane) * 10 // (7 * k))
        out.append(make_scene(seed * 100 + s, k, per, 0.3, 0.1, 200.0, plane_points=plane))
    return out


def height_field_mesh(n: int = 24, seed: int = 3):
    """An open, asymmetric triangle mesh z = f(x, y) over the unit square:
    an n x n vertex grid, two counter-clockwise triangles per cell (normals
    towards +z).  Returns (points n^2 x 3 float32, faces 2(n-1)^2 x 3 int32)."""
    rng = _rng(seed)
    g = np.linspace(0.0, 1.0, n, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="ij")
    Z = 0.2 * np.sin(3 * X) * np.cos(2 * Y) + 0.1 * X * X + 0.003 * rng.standard_normal(X.shape)
    P = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).ravel()
    b, c, d = a + n, a + n + 1, a + 1
    F = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    return P, F.astype(np.int32)


def symmetries() -> list:
    """Rigid self-maps of the model (identity and the 180 degree turn about
    the cylinder axis x=50, y=30)."""
    S = np.eye(4)
    S[0, 0] = S[1, 1] = -1
    S[0, 3] = 100.0
    S[1, 3] = 60.0
    return [np.eye(4), S]


def rotation_angle_deg(dR) -> float:
    """Rotation angle of a 3x3 matrix, atan2(|skew(dR)|, (tr dR - 1)/2):
    well conditioned near 0 and pi, unlike acos of the trace."""
    dR = np.asarray(dR, np.float64)
    s = 0.5 * math.sqrt((dR[2, 1] - dR[1, 2]) ** 2 + (dR[0, 2] - dR[2, 0]) ** 2 +
                        (dR[1, 0] - dR[0, 1]) ** 2)
    c = 0.5 * (np.trace(dR) - 1.0)
    return math.degrees(math.atan2(s, c))


def pose_error(T, gt_poses, diameter: float):
    """(instance, rotation error deg, translation error mm) of the closest
    ground-truth pose, modulo the model's symmetries."""
    T = np.asarray(T, np.float64)
    if T.shape == (3, 4):
        T = np.vstack([T, [0, 0, 0, 1]])
    best = (None, 1e9, 1e9)
    for k, G in enumerate(gt_poses):
        for S in symmetries():
            GS = G @ S
            ang = rotation_angle_deg(GS[:3, :3].T @ T[:3, :3])
            dt = float(np.linalg.norm(GS[:3, 3] - T[:3, 3]))
            if ang / 0.1 + dt / (0.001 * diameter) < best[1] / 0.1 + best[2] / (0.001 * diameter):
                best = (k, ang, dt)
    return best
