"""MLS normals: estimateNormalMLS (util.cpp:336-457).

The oracle (orc_estimate_normals_mls) is pinned by
  * its exp against numpy's float64 exp (below one ulp) and IEEE edge cases,
  * Eigen's unitOrthogonal rule (orthonormality, the two branches' closed
    forms),
  * closed forms of the fit: a plane is reproduced (points stay, normal =
    the plane's), the apex of a paraboloid is a fixed point of the order-2
    fit with normal +z, an order-1 fit at the apex lifts the point by the
    weighted mean height,
  * a sphere: the fit pulls noisy points towards the sphere and turns the
    normals radial,
  * an independent float64 numpy formulation (eigh for the PCA, lstsq on
    sqrt-weighted rows for the fit, the normal from the fitted surface's
    gradient) on noisy surfaces,
and the GPU path is compared with the oracle bit for bit.
"""
import math

import numpy as np
import pytest

from conftest import rng


def f3(*v):
    return np.array(v, np.float32)


# ------------------------------------------------------------------ helpers

def surface_cloud(n, seed, noise=0.002, kind="wave"):
    """points on z = f(x, y) over [-1, 1]^2 with Gaussian noise along z"""
    r = rng(seed)
    xy = r.uniform(-1, 1, (n, 2))
    x, y = xy[:, 0], xy[:, 1]
    if kind == "wave":
        z = 0.2 * np.sin(2 * x) * np.cos(1.5 * y) + 0.05 * x * y
    else:
        z = 0.1 * x - 0.2 * y
    z = z + noise * r.standard_normal(n)
    return np.stack([x, y, z], 1).astype(np.float32)


def numpy_mls(pts, q, radius, order, vp=None, invert=False):
    """estimateNormalMLS for one query in float64: returns (point, normal)
    or None when it keeps its inputs"""
    P = pts.astype(np.float64)
    p = P[q]
    d2 = ((pts - pts[q]) ** 2).sum(1)
    nb = P[d2 < np.float32(radius) ** 2]
    if len(nb) < 3:
        return None
    c = nb.mean(0)
    w, V = np.linalg.eigh((nb - c).T @ (nb - c))
    n = V[:, np.argmin(np.abs(w))]
    mean = p - (p @ n - n @ c) * n
    # Darboux frame: v orthogonal to n as Eigen's unitOrthogonal picks it
    if abs(n[0]) > 1e-5 * abs(n[2]) or abs(n[1]) > 1e-5 * abs(n[2]):
        v = np.array([-n[1], n[0], 0.0]) / math.hypot(n[0], n[1])
    else:
        v = np.array([0.0, -n[2], n[1]]) / math.hypot(n[1], n[2])
    u = np.cross(n, v)
    e = nb - mean
    wt = np.exp(-(e ** 2).sum(1) / np.float32(radius) ** 2)
    uc, vc, h = e @ u, e @ v, e @ n
    powers = [(i, j) for i in range(order + 1) for j in range(order + 1 - i)]
    X = np.stack([uc ** i * vc ** j for i, j in powers], 1)
    sw = np.sqrt(wt)
    coef, *_ = np.linalg.lstsq(X * sw[:, None], h * sw, rcond=None)
    cof = dict(zip(powers, coef))
    point = mean + cof[(0, 0)] * n
    normal = n - cof[(1, 0)] * u - cof[(0, 1)] * v    # surface gradient at (0, 0)
    normal /= np.linalg.norm(normal)
    if vp is not None and np.all(np.isfinite(vp)):
        val = normal @ (np.asarray(vp, np.float64) - point)
    else:
        val = normal[2]
    if (val > 0) if invert else (val < 0):
        normal = -normal
    return point, normal


def angle_deg(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    c = np.clip((a * b).sum(-1) / (np.linalg.norm(a, axis=-1) * np.linalg.norm(b, axis=-1)),
                -1, 1)
    return np.degrees(np.arccos(c))


# ------------------------------------------------------------------ CPU

def test_expf_against_numpy(oracle_mod, H):
    L = oracle_mod.lib()
    xs = np.concatenate([np.linspace(-103, 88.7, 40001), np.linspace(-8, 0, 20001),
                         rng(1).uniform(-20, 5, 5000)]).astype(np.float32)
    got = np.array([L.orc_expf(float(x)) for x in xs], np.float32)
    host = np.array([H.ch_expf(float(x)) for x in xs], np.float32)
    assert np.array_equal(got.view(np.uint32), host.view(np.uint32))
    ref = np.exp(xs.astype(np.float64))
    ok = (ref > 1.2e-38) & (ref < 3.4e38)
    ulp = np.abs(got[ok] - ref[ok]) / np.spacing(ref[ok].astype(np.float32))
    assert ulp.max() < 1.0
    assert L.orc_expf(0.0) == 1.0 and H.ch_expf(0.0) == 1.0
    assert L.orc_expf(-200.0) == 0.0 and math.isinf(L.orc_expf(100.0))
    assert math.isnan(L.orc_expf(float("nan"))) and math.isnan(H.ch_expf(float("nan")))
    assert L.orc_expf(float("-inf")) == 0.0 and math.isinf(L.orc_expf(float("inf")))


def test_unit_orthogonal(oracle_mod, H):
    L = oracle_mod.lib()
    o = np.zeros(3, np.float32)
    L.orc_unit_orthogonal(f3(0, 0, 1), o)
    assert np.array_equal(o, f3(0, -1, 0))
    L.orc_unit_orthogonal(f3(1, 0, 0), o)
    assert np.array_equal(np.abs(o), f3(0, 1, 0))
    L.orc_unit_orthogonal(f3(3e-6, -2e-6, 1), o)        # both below 1e-5 |z|: (0, -z, y)
    assert o[0] == 0
    L.orc_unit_orthogonal(f3(2e-5, 0, 1), o)            # x above it: (-y, x, 0)
    assert o[2] == 0 and o[1] == 1
    r = rng(2)
    for _ in range(2000):
        n = r.standard_normal(3)
        if r.random() < 0.3:
            n[:2] *= 10.0 ** r.uniform(-8, -4)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        L.orc_unit_orthogonal(n, o)
        h = np.zeros(3, np.float32)
        H.ch_unit_orthogonal(n, h)
        assert np.array_equal(o.view(np.uint32), h.view(np.uint32))
        assert abs(float(o.astype(np.float64) @ n)) < 1e-6
        assert abs(np.linalg.norm(o) - 1) < 1e-6


def test_mls_plane_is_fixed(oracle_mod):
    P = surface_cloud(3000, 3, noise=0.0, kind="plane")
    q = np.flatnonzero((np.abs(P[:, :2]) < 0.6).all(1))[:300]
    nrm = np.full_like(P, np.nan)
    want = f3(-0.1, 0.2, 1) / np.linalg.norm(f3(-0.1, 0.2, 1))
    for order in (1, 2, 3):
        op, on, tot = oracle_mod.estimate_normals_mls(P, nrm, np.arange(len(P)), q, 0.2, order)
        assert tot > 10 * len(q)
        assert np.abs(op[q] - P[q]).max() < 2e-6
        assert angle_deg(on[q], want).max() < 2e-3, order
        untouched = np.setdiff1d(np.arange(len(P)), q)
        assert np.array_equal(op[untouched], P[untouched])
        assert np.isnan(on[untouched]).all()


def test_mls_paraboloid_apex(oracle_mod):
    # symmetric grid on z = a x^2 + b y^2: the apex's PCA normal is +-z; an
    # order-2 fit reproduces the surface, so the apex stays with normal +z;
    # an order-1 fit lifts it to the weighted mean height
    a, b, h = 0.8, 0.3, 0.02
    k = np.arange(-12, 13)
    X, Y = np.meshgrid(k * h, k * h, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), (a * X * X + b * Y * Y).ravel()], 1).astype(np.float32)
    apex = int(np.flatnonzero((P[:, 0] == 0) & (P[:, 1] == 0))[0])
    nrm = np.full_like(P, np.nan)
    R = 0.15
    op, on, _ = oracle_mod.estimate_normals_mls(P, nrm, np.arange(len(P)), [apex], R, 2)
    assert np.abs(op[apex]).max() < 1e-6
    assert angle_deg(on[apex], f3(0, 0, 1)) < 1e-3
    op1, on1, _ = oracle_mod.estimate_normals_mls(P, nrm, np.arange(len(P)), [apex], R, 1)
    nb = P[((P - P[apex]) ** 2).sum(1) < np.float32(R) ** 2].astype(np.float64)
    zbar = nb[:, 2].mean()                   # the PCA plane passes through the centroid
    e = nb - np.array([0, 0, zbar])
    w = np.exp(-(e ** 2).sum(1) / np.float32(R) ** 2)
    lift = zbar + (w * e[:, 2]).sum() / w.sum()
    assert abs(op1[apex, 2] - lift) < 1e-6 and abs(op1[apex, :2]).max() < 1e-6
    assert angle_deg(on1[apex], f3(0, 0, 1)) < 1e-3


def test_mls_sphere_denoise(oracle_mod):
    r = rng(4)
    n, Rs, sigma = 20000, 1.0, 0.01
    d = r.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    P = (d * (Rs + sigma * r.standard_normal((n, 1)))).astype(np.float32)
    q = np.arange(0, n, 10)
    nrm = np.full_like(P, np.nan)
    # view point at the centre, inverted: normals point away from it
    op, on, _ = oracle_mod.estimate_normals_mls(P, nrm, np.arange(n), q, 0.15, 2, invert=1,
                                                vp=f3(0, 0, 0))
    before = np.abs(np.linalg.norm(P[q], axis=1) - Rs)
    after = np.abs(np.linalg.norm(op[q], axis=1) - Rs)
    assert np.sqrt((after ** 2).mean()) < 0.4 * np.sqrt((before ** 2).mean())
    assert np.percentile(angle_deg(on[q], op[q]), 99) < 3.0


@pytest.mark.parametrize("order", [1, 2, 3])
def test_mls_matches_float64_formulation(oracle_mod, order):
    P = surface_cloud(4000, 10 + order, noise=0.003)
    q = np.flatnonzero((np.abs(P[:, :2]) < 0.8).all(1))[:150]
    nrm = np.full_like(P, np.nan)
    vp = f3(0.3, -0.2, 5.0)
    op, on, _ = oracle_mod.estimate_normals_mls(P, nrm, np.arange(len(P)), q, 0.12, order,
                                                vp=vp)
    for t in q:
        ref = numpy_mls(P, t, 0.12, order, vp=vp)
        assert ref is not None
        assert np.abs(op[t] - ref[0]).max() < 2e-5, (order, t)
        assert angle_deg(on[t], ref[1]) < 0.02, (order, t)


def test_mls_skips_and_errors(oracle_mod):
    P = surface_cloud(500, 5)
    P[7] = np.nan
    lone = len(P)
    P = np.vstack([P, f3(50, 50, 50)])               # no neighbours but itself
    nrm = np.tile(f3(0, 0, 1), (len(P), 1))
    valid = np.flatnonzero(np.isfinite(P).all(1))
    op, on, _ = oracle_mod.estimate_normals_mls(P, nrm, valid, [7, lone], 0.2, 2)
    assert np.array_equal(op[lone], P[lone]) and np.array_equal(on[lone], nrm[lone])
    assert np.isnan(op[7]).all() and np.array_equal(on[7], nrm[7])
    for bad in (0, 4):
        with pytest.raises(ValueError):
            oracle_mod.estimate_normals_mls(P, nrm, valid, [0], 0.2, bad)
    with pytest.raises(ValueError):
        oracle_mod.estimate_normals_mls(P, nrm, valid, [0], 0.0, 2)


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("order", [1, 2, 3])
def test_gpu_mls_bitwise(order, oracle_mod):
    import surface_match_amd as sm
    P = surface_cloud(30000, 20 + order, noise=0.004)
    P[::997] = np.nan
    valid = np.flatnonzero(np.isfinite(P).all(1))
    q = np.arange(0, len(P), 3)
    for vp, invert in ((f3(0.5, 0.5, 4), 0), (f3(np.nan, np.nan, np.nan), 0),
                       (f3(0, 0, -4), 1)):
        pc = sm.PointCloud(P.copy())
        pc.viewPoint = vp
        st = sm.estimateNormalMLS(pc, q, 0.06, order, bool(invert), validIndices=valid)
        op, on, tot = oracle_mod.estimate_normals_mls(P, np.full_like(P, np.nan), valid, q, 0.06,
                                                      order, invert=invert, vp=vp)
        print("mls order", order, "vp", vp, "invert", invert, "neighbours", tot,
              "device ms", round(st["device_ms"], 3), flush=True)
        assert st["neighbours"] == tot
        assert np.array_equal(pc.point.view(np.uint32), op.view(np.uint32))
        assert np.array_equal(pc.normal.view(np.uint32), on.view(np.uint32))


@pytest.mark.gpu
def test_gpu_mls_errors_and_skips(oracle_mod):
    import surface_match_amd as sm
    P = surface_cloud(2000, 7)
    pc = sm.PointCloud(P.copy())
    for bad in (0, 4):
        with pytest.raises(ValueError, match="order"):
            sm.estimateNormalMLS(pc, [0], 0.1, bad)
    with pytest.raises(ValueError, match="radius"):
        sm.estimateNormalMLS(pc, [0], -1.0, 2)
    with pytest.raises(ValueError, match="query"):
        sm.estimateNormalMLS(pc, [len(P)], 0.1, 2)
    # a tiny radius: every query keeps its point and (absent) normal
    pc = sm.PointCloud(P.copy())
    st = sm.estimateNormalMLS(pc, np.arange(len(P)), 1e-6, 2)
    assert st["neighbours"] == len(P)
    assert np.array_equal(pc.point, P) and np.isnan(pc.normal).all()
    # no queries
    pc = sm.PointCloud(P.copy())
    assert sm.estimateNormalMLS(pc, [], 0.1, 2)["neighbours"] == 0


@pytest.mark.gpu
def test_gpu_mls_sphere_1m(oracle_mod):
    """1M points on a noisy sphere: the GPU fit denoises and turns the normals
    radial, and a sample of queries equals the oracle's bit for bit"""
    import surface_match_amd as sm
    r = rng(8)
    n, sigma = 1_000_000, 0.004
    d = r.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    P = (d * (1.0 + sigma * r.standard_normal((n, 1)))).astype(np.float32)
    q = np.arange(0, n, 4)
    pc = sm.PointCloud(P.copy())
    pc.viewPoint = f3(0, 0, 0)
    st = sm.estimateNormalMLS(pc, q, 0.03, 2, invert=True)
    print("sphere 1M:", st, flush=True)
    before = np.abs(np.linalg.norm(P[q], axis=1) - 1.0)
    after = np.abs(np.linalg.norm(pc.point[q], axis=1) - 1.0)
    assert np.sqrt((after ** 2).mean()) < 0.5 * np.sqrt((before ** 2).mean())
    ang = angle_deg(pc.normal[q], pc.point[q])       # radial = outward normal
    assert np.median(ang) < 1.5 and np.percentile(ang, 99) < 6.0
    sub = q[::500]
    op, on, _ = oracle_mod.estimate_normals_mls(P, np.full_like(P, np.nan), np.arange(n), sub,
                                                0.03, 2, invert=1, vp=f3(0, 0, 0))
    assert np.array_equal(pc.point[sub].view(np.uint32), op[sub].view(np.uint32))
    assert np.array_equal(pc.normal[sub].view(np.uint32), on[sub].view(np.uint32))
