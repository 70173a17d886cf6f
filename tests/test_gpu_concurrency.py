"""Concurrent calls on one Detector.  The reference's matchScene only reads
the trained model (ppf.cpp:149-510), so several threads may match scenes on
one Detector at once; here every call takes its own context (stream,
scratch, scene copy) from the detector's pool and reads the model under a
shared lock.  Threads matching different scenes, and threads matching one
resident scene (serialised by the scene), must return exactly their serial
results."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scenes():
    from surface_match_amd import synth
    w = synth.workload(1)
    out = [w["scene"].points]
    for k in range(3):
        sc = synth.make_scene(0x5EED0100 + k, 2 + k, 20000, 0.3, 0.1, 200.0, plane_points=30000)
        out.append(sc.points)
    return w["model"], out


def _run_threads(fns):
    res, err = [None] * len(fns), []

    def body(i):
        try:
            res[i] = fns[i]()
        except Exception as e:  # noqa: BLE001 -- reported below
            err.append(e)
    th = [threading.Thread(target=body, args=(i,)) for i in range(len(fns))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not err, err
    return res


def test_concurrent_matches_equal_serial():
    from surface_match_amd import Detector, MatchParam, PointCloud, ResidentScene, synth
    (mp, mn), scenes = _scenes()
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    param = MatchParam(numMatches=3)
    clouds = []
    for p in scenes:
        c = PointCloud(p)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        clouds.append(c)
    serial = [det.matchScene(c, 0.04, 0.1, 0.2, param) for c in clouds]
    assert all(len(s[1]) > 0 for s in serial)

    # host-array matchScene: four threads, each its own scene, three rounds
    def host(i):
        return [det.matchScene(clouds[i], 0.04, 0.1, 0.2, param) for _ in range(3)]
    for i, runs in enumerate(_run_threads([lambda i=i: host(i) for i in range(4)])):
        for P, S in runs:
            assert np.array_equal(P, serial[i][0]) and np.array_equal(S, serial[i][1]), i

    # resident scenes: each thread its own scene, plus two threads sharing one
    rs = [ResidentScene(det, c) for c in clouds]
    fns = [lambda i=i: [rs[i].match(0.04, 0.1, 0.2, param)[:2] for _ in range(2)]
           for i in range(4)]
    fns.append(lambda: [rs[0].match(0.04, 0.1, 0.2, param)[:2] for _ in range(2)])
    out = _run_threads(fns)
    for i, runs in enumerate(out):
        j = i if i < 4 else 0
        for P, S in runs:
            assert np.array_equal(P, serial[j][0]) and np.array_equal(S, serial[j][1]), i
    print(f"{len(clouds)} scenes: concurrent host-array and resident matches equal serial results")


def test_batch_with_repeated_scene_equals_single():
    """A batch naming one scene twice matches it once and reports its result
    at both entries."""
    from surface_match_amd import Detector, MatchParam, PointCloud, ResidentScene, synth
    (mp, mn), scenes = _scenes()
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    param = MatchParam(numMatches=3)
    rs = []
    for p in scenes[:2]:
        c = PointCloud(p)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        rs.append(ResidentScene(det, c))
    single = [r.match(0.04, 0.1, 0.2, param)[:2] for r in rs]
    many, _ = ResidentScene.match_many([rs[0], rs[1], rs[0]], 0.04, 0.1, 0.2, param)
    for (P, S), j in zip(many, [0, 1, 0]):
        assert np.array_equal(P, single[j][0]) and np.array_equal(S, single[j][1])
