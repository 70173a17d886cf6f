"""The scene upload on the GPU: removeNan (util.cpp:86-105) and
computeBoundingBox (util.cpp:178-207) as matchScene applies them to the
caller's cloud (ppf.cpp:153,170-176), against the oracle's restatement.

Clouds carry NaN and infinite coordinates and normals, and extremes held by
several points, among them +0 and -0 (the reference keeps the earliest point
holding the extreme: a later equal value, of either sign, replaces nothing).
Sizes cover one point, a partial wave, a partial block and 4M points.  The
valid indices and the box bits must equal the oracle's; the points and
normals must come back as given.
"""
import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu


def _cloud(n, seed, normals):
    r = rng(seed)
    p = r.uniform(-50, 50, (n, 3)).astype(np.float32)
    p[:, 0] = np.abs(p[:, 0])              # min x is a zero of either sign
    p[:, 1] = -np.abs(p[:, 1])             # max y is a zero of either sign
    k = max(1, n // 50)
    zx = r.choice(n, k)
    p[zx, 0] = np.where(r.random(k) < 0.5, np.float32(0.0), np.float32(-0.0))
    zy = r.choice(n, k)
    p[zy, 1] = np.where(r.random(k) < 0.5, np.float32(0.0), np.float32(-0.0))
    p[r.choice(n, k), 2] = np.float32(-50.0)  # a repeated minimum
    p[r.choice(n, k), 2] = np.float32(50.0)   # a repeated maximum
    bad = r.choice(n, k)
    p[bad, r.integers(0, 3, k)] = np.where(r.random(k) < 0.5, np.nan, np.inf).astype(np.float32)
    nrm = None
    if normals:
        nrm = r.normal(size=(n, 3)).astype(np.float32)
        nrm[r.choice(n, k), r.integers(0, 3, k)] = np.nan
    return p, nrm


@pytest.mark.parametrize("n", [1, 63, 64, 65, 257, 100_000, 4_000_000])
@pytest.mark.parametrize("normals", [False, True])
def test_upload_valid_indices_and_box(n, normals, oracle_mod):
    from surface_match_amd import Detector, PointCloud, ResidentScene
    O = oracle_mod
    p, nrm = _cloud(n, 7 + n, normals)
    det = Detector()
    rs = ResidentScene(det, PointCloud(p, nrm) if normals else PointCloud(p))
    gp, gn, valid, box = rs.download()
    assert np.array_equal(gp.view(np.uint32), p.view(np.uint32))
    if normals:
        assert np.array_equal(gn.view(np.uint32), nrm.view(np.uint32))
    else:
        assert gn is None
    ov = O.remove_nan(p, nrm)
    assert np.array_equal(valid, ov)
    if len(ov):
        ob = O.bounding_box(p, ov)
        assert np.array_equal(box.view(np.uint32), ob.view(np.uint32)), (box, ob)
    else:
        # no valid point: the box stays zero (no later stage reads it)
        assert not box.any()
    print(f"n={n} normals={normals}: {len(ov)} valid, box {box}")


def test_upload_no_valid_point_and_caller_box(oracle_mod):
    from surface_match_amd import BoundingBox, Detector, PointCloud, ResidentScene
    p = np.full((300, 3), np.nan, np.float32)
    rs = ResidentScene(Detector(), PointCloud(p))
    _, _, valid, box = rs.download()
    assert len(valid) == 0 and not box.any()
    q, _ = _cloud(1000, 3, False)
    c = PointCloud(q)
    c.box = BoundingBox(np.array([-1, -2, -3], np.float32), np.array([4, 5, 6], np.float32))
    _, _, valid, box = ResidentScene(Detector(), c).download()
    assert np.array_equal(valid, oracle_mod.remove_nan(q))
    assert np.array_equal(box, [-1, -2, -3, 4, 5, 6])
