"""GPU parity at the benchmark's sizes against committed oracle fixtures.

* The 1M-point scene (100k-point box+cylinder model; 10 instances, 40%
  occlusion, sigma 0.2 mm, 20% clutter): matchScene's results against the
  oracle's, and the vote accumulators and candidates of 32 sampled key
  points of the GPU's own sampled scene bit for bit (ppf.cpp:262-391).
* 64 scenes of uneven size (log-uniform 50k-500k points, 1-8 instances)
  matched in one batch: every scene's results against the oracle's match of
  that scene alone.
* The batch's own pose hypotheses (vote + clustering, at least 512 of them)
  refined by the batched point-to-plane refiner, sampled ones against the
  FP64 oracle refinement (icp.cpp:214-278 with the point-to-plane solve).

Fixtures: tests/golden/make_match_oracle.py.  When this run's generator
produces other clouds than the fixture's (SHA-256), the oracle runs here.
"""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import rng
import parity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def workload2():
    from surface_match_amd import synth
    return synth.workload(2)


def test_workload2_matches_oracle(workload2, oracle_mod):
    from surface_match_amd import Detector, PointCloud, MatchParam, MatchResult, ResidentScene, synth
    O = oracle_mod
    w = workload2
    mp, mn = w["model"]
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    info = det.model_info()
    g = parity.load_fixture("workload2_oracle.json")
    nm = g["match"]["numMatches"]
    if (g["scene_sha256"] == parity.digest(w["scene"].points) and
            g["model_sha256"] == parity.digest(mp)):
        OP, OS = parity.fixture_result(g)
        src = "fixture"
    else:
        OP, OS = parity.oracle_match(O, O.train(mp, mn, 0.04), w["scene"].points, nm)
        src = "computed"
    cloud = PointCloud(w["scene"].points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    param = MatchParam(numMatches=nm)
    rs = ResidentScene(det, cloud)
    P, S, st = rs.match(0.04, 0.1, 0.2, param)
    print(f"oracle ({src}) scores {OS}; gpu scores {S}; {st['nSampled']} sampled, "
          f"{st['nKeypoints']} key points, {st['nPairs']} pairs, {st['nVotes']} votes")
    at_oracle = parity.oracle_pose_scores(rs, OP, info.nReSampled)
    a, t, s = parity.check_results(P, S, OP, OS, info.diameter, at_oracle)
    print(f"{len(S)} results: max {a:.2e} deg, {t:.2e} mm from the oracle; own-pose score "
          f"max |diff| {s:.2e} (one inlier = {1 / info.nReSampled:.2e})")
    # the host-array call (the drop-in matchScene) returns the same results
    mr = MatchResult()
    P2, S2 = det.matchScene(cloud, 0.04, 0.1, 0.2, param, matchResult=mr)
    assert np.array_equal(P2, P) and np.array_equal(S2, S)
    voted, n = parity.check_votes(det, O, mr.sampledScene.point, mr.sampledScene.normal, 32, 202)
    print(f"{voted} of {n} sampled key points voted, accumulators and candidates bit-exact")
    assert voted >= 24


@pytest.fixture(scope="module")
def batch64():
    """The 20k-point model trained on the GPU, the 64 scenes resident, the
    batched match and the fixture."""
    from surface_match_amd import Detector, PointCloud, MatchParam, ResidentScene, synth
    g = parity.load_fixture("batch64_oracle.json")
    mp, mn = synth.make_model(20000, 0x5EED0001)
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    scenes = synth.batch_scenes(0x5EED0003, 64)
    rs = []
    for sc in scenes:
        c = PointCloud(sc.points)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        rs.append(ResidentScene(det, c))
    param = MatchParam(numMatches=g["match"]["numMatches"])
    many, st = ResidentScene.match_many(rs, 0.04, 0.1, 0.2, param)
    return dict(g=g, model=(mp, mn), det=det, scenes=scenes, rs=rs, many=many, st=st,
                param=param)


def test_batch64_matches_oracle(batch64, oracle_mod):
    from surface_match_amd import synth
    O = oracle_mod
    b = batch64
    g, det = b["g"], b["det"]
    info = det.model_info()
    same_model = g["model_sha256"] == parity.digest(b["model"][0])
    om = None
    worst = [0.0, 0.0, 0.0]
    n_res = 0
    for i, (sc, rs, (P, S), rec) in enumerate(zip(b["scenes"], b["rs"], b["many"], g["scenes"])):
        if same_model and rec["scene_sha256"] == parity.digest(sc.points):
            OP, OS = parity.fixture_result(rec)
        else:
            om = om or O.train(*b["model"], 0.04)
            OP, OS = parity.oracle_match(O, om, sc.points, g["match"]["numMatches"])
        at_oracle = parity.oracle_pose_scores(rs, OP, info.nReSampled)
        try:
            w = parity.check_results(P, S, OP, OS, info.diameter, at_oracle)
        except AssertionError as e:
            raise AssertionError(f"scene {i} ({len(sc.points)} points): {e}") from None
        worst = [max(x, y) for x, y in zip(worst, w)]
        n_res += len(S)
    st = b["st"]
    print(f"64 scenes, {sum(len(s.points) for s in b['scenes'])} points, {n_res} results "
          f"equal to the oracle's scene by scene (max {worst[0]:.2e} deg, {worst[1]:.2e} mm; "
          f"own-pose score max |diff| {worst[2]:.2e}); {st['nSparseIcp']} sparse and "
          f"{st['nDenseIcp']} dense refinements, batch {st['msTotal']:.1f} ms")


def test_batch_hypotheses_refined_vs_fp64_oracle(batch64, oracle_mod):
    """The clustered hypotheses of the 64-scene batch (refinement off, so the
    batch returns every hypothesis with votes >= the vote threshold; key
    point fraction 0.2 gives at least 512 of them), each refined by the
    batched point-to-plane refiner (10 iterations) against its scene's
    sampled cloud with the normals the match estimated; sampled hypotheses
    against the FP64 oracle refinement of the same start on the same
    clouds."""
    from surface_match_amd import MatchParam, MatchResult, PointCloud, ResidentScene, synth
    from surface_match_amd.icp import IcpTarget
    O = oracle_mod
    b = batch64
    det = b["det"]
    info = det.model_info()
    D = float(info.diameter)
    sp, sn, _, _ = det.sampled_model()
    param = MatchParam(numMatches=100000, sparsePoseRefinement=0, densePoseRefinement=0)
    hyps, _ = ResidentScene.match_many(b["rs"], 0.04, 0.2, 0.2, param)
    total = sum(len(P) for P, _ in hyps)
    print(f"{total} hypotheses from the batch vote")
    assert total >= 512
    crit = dict(iterations=10, reject_dist=0.1 * D, mse_min=0.0, mse_max=1e30, tolerance=0.0,
                mode=1, scoring_dist=0.01 * D)
    r = rng(303)
    flat = [(s, h) for s, (P, _) in enumerate(hyps) for h in range(len(P))]
    picks = set(tuple(flat[k]) for k in r.choice(len(flat), 24, replace=False))
    refined = 0
    errs = []
    for s, (P, _) in enumerate(hyps):
        if len(P) == 0:
            continue
        sc = b["scenes"][s]
        cloud = PointCloud(sc.points)
        cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        mr = MatchResult()
        det.matchScene(cloud, 0.04, 0.2, 0.2, param, matchResult=mr)
        dst, dstn = mr.sampledScene.point, mr.sampledScene.normal
        tgt = IcpTarget(dst, dstn, 0.08 * D)
        poses, res, _ = tgt.refine(sp, sn, P, batched=True, **crit)
        refined += len(P)
        mine = [h for (ss, h) in picks if ss == s]
        if not mine:
            continue
        grid = O.OracleGrid(dst, 0.08 * D)
        for h in mine:
            c = O.Criteria(10, 0.1 * D, 0.0, 1e30, 0.0)
            ores = O.IcpResult()
            O.lib().orc_icp_regist(len(sp), sp, sn, C.byref(grid._g), dstn,
                                   np.ascontiguousarray(P[h][:3].ravel(), np.float32),
                                   C.byref(c), 1, C.byref(ores))
            OPh = np.array(ores.pose, np.float64).reshape(3, 4)
            a = synth.rotation_angle_deg(OPh[:, :3].T @ poses[h][:3, :3])
            dt = float(np.linalg.norm(OPh[:, 3] - poses[h][:3, 3]))
            errs.append((s, h, a, dt))
    assert refined == total
    bad = [e for e in errs if not (e[2] < 0.1 and e[3] < 0.001 * D)]
    print(f"{refined} hypotheses refined; {len(errs)} sampled against the FP64 oracle: max "
          f"{max(e[2] for e in errs):.2e} deg, {max(e[3] for e in errs):.2e} mm; outside "
          f"0.1 deg / 0.1% D: {bad}")
    assert not bad
