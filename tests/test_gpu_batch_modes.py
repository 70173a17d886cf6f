"""ppf_match_scenes_resident against ppf_match_scene_resident scene by scene
under every refinement setting (sparse and dense, sparse only, dense only,
none), for a single scene and for three: equal bit for bit, and two batched
calls in a row equal.  The batched match clusters each scene on its lanes
after a vote whose pose buffer may grow, and prepares the refinement beside
the vote; each setting exercises a different part of that."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    from surface_match_amd import Detector, PointCloud, ResidentScene, synth
    w = synth.workload(1)
    det = Detector()
    det.trainModel(PointCloud(*w["model"]), 0.04)
    rs = []
    for k, pts in enumerate([w["scene"].points] +
                            [synth.make_scene(0x5EED0200 + k, 2 + k, 20000, 0.3, 0.1, 200.0,
                                              plane_points=30000).points for k in range(2)]):
        c = PointCloud(pts)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        rs.append(ResidentScene(det, c))
    return det, rs


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


@pytest.mark.parametrize("mode", ["both", "sparse", "dense", "none"])
def test_batch_equals_single(scenes, mode):
    from surface_match_amd import MatchParam, ResidentScene
    _, rs = scenes
    kw = {"both": {}, "sparse": dict(densePoseRefinement=0),
          "dense": dict(sparsePoseRefinement=0),
          "none": dict(sparsePoseRefinement=0, densePoseRefinement=0)}[mode]
    param = MatchParam(numMatches=5, **kw)
    for group in ([rs[0]], rs):
        single = [r.match(0.04, 0.1, 0.2, param)[:2] for r in group]
        m1, _ = ResidentScene.match_many(group, 0.04, 0.1, 0.2, param)
        m2, _ = ResidentScene.match_many(group, 0.04, 0.1, 0.2, param)
        assert _same(m1, single), (mode, len(group))
        assert _same(m2, m1), (mode, len(group))
        print(f"{mode}, {len(group)} scene(s): results {[len(S) for _, S in m1]}")


@pytest.mark.parametrize("kw", [dict(), dict(smoothNormal=False, invertNormal=True),
                                dict(knnNormal=20)])
def test_batch_mixes_scenes_with_and_without_normals(scenes, kw):
    """A batch whose middle scene brings its own normals (random directions,
    NaN on 1 % of its points) between two whose normals are estimated:
    every scene's result equals its single-scene match, with smoothing,
    without smoothing and inverted, and with 20 neighbours (the estimation
    that keeps up to 32 neighbours a point)."""
    from surface_match_amd import MatchParam, PointCloud, ResidentScene, synth
    det, rs = scenes
    pts = synth.make_scene(0x5EED0210, 3, 20000, 0.3, 0.1, 200.0, plane_points=30000).points
    r = np.random.default_rng(21)
    nrm = r.normal(size=pts.shape).astype(np.float32)
    nrm[r.choice(len(nrm), len(nrm) // 100, replace=False)] = np.nan
    c = PointCloud(pts, nrm)
    c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    group = [rs[1], ResidentScene(det, c), rs[2]]
    param = MatchParam(numMatches=5, **kw)
    single = [x.match(0.04, 0.1, 0.2, param)[:2] for x in group]
    many, _ = ResidentScene.match_many(group, 0.04, 0.1, 0.2, param)
    assert _same(many, single), kw
    print(f"{kw}: results {[len(S) for _, S in many]}")


def test_batch_with_a_scene_too_large_for_shared_grids(scenes):
    """A scene with one point 10 m away: its grids (normals, vote,
    refinement) exceed the dense cell limit, so the batch builds them scene
    by scene and estimates every scene's normals one by one (the fallbacks
    of the shared builds).  Every scene's result equals its single-scene
    match."""
    from surface_match_amd import MatchParam, PointCloud, ResidentScene, synth
    det, rs = scenes
    pts = synth.make_scene(0x5EED0220, 2, 20000, 0.3, 0.1, 200.0, plane_points=30000).points
    pts = np.vstack([pts, np.array([[1.0e4, 1.0e4, 1.0e4]], np.float32)])
    c = PointCloud(pts)
    c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    group = [rs[1], ResidentScene(det, c), rs[2]]
    param = MatchParam(numMatches=5)
    single = [x.match(0.04, 0.1, 0.2, param)[:2] for x in group]
    many, _ = ResidentScene.match_many(group, 0.04, 0.1, 0.2, param)
    assert _same(many, single)
    assert len(many[1][1]) > 0
    print(f"results {[len(S) for _, S in many]}")
