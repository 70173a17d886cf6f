"""Checks shared by the full-size GPU parity tests: a match against the
oracle's result kept as a fixture (or recomputed when this run's generator
produced other clouds), and the vote accumulators of sampled key points of
the GPU's own sampled scene against the oracle's vote loop on the same
inputs (ppf.cpp:262-391)."""
import hashlib
import json
import math
import os

import numpy as np

from conftest import rng

HERE = os.path.dirname(os.path.abspath(__file__))
TWO_PI_F = np.float32(2.0) * np.float32(math.pi)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, np.float32).tobytes()).hexdigest()


def load_fixture(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def fixture_result(rec):
    S = np.array(rec["scores_f32_bits"], np.uint32).view(np.float32)
    P = np.array(rec["poses_3x4"], np.float32).reshape(-1, 3, 4)
    return P, S


def oracle_match(O, model, points, num_matches, threads=None):
    from surface_match_amd import synth
    P, S, _ = O.match(model, points, None, 0.04, 0.1, 0.2,
                      param=O.default_match_param(numMatches=num_matches),
                      vp=np.array(synth.VIEWPOINT, np.float32),
                      threads=threads or min(os.cpu_count() or 1, 16))
    return P, S


def check_results(P, S, OP, OS, diameter, at_oracle_scores):
    """Same count and order as the oracle; every pose within 0.1 deg and
    0.1% D of the oracle's; the GPU's score evaluated at the oracle's pose
    equal to the oracle's score within 1e-5.  Returns the largest
    differences (rotation deg, translation mm, own-pose score)."""
    from surface_match_amd import synth
    assert len(S) == len(OS), (S, OS)
    worst_a = worst_t = 0.0
    for T, Q in zip(P, OP):
        Q4 = np.vstack([Q.astype(np.float64), [0, 0, 0, 1]])
        ang = synth.rotation_angle_deg(Q4[:3, :3].T @ T[:3, :3])
        dt = float(np.linalg.norm(Q4[:3, 3] - T[:3, 3]))
        assert ang < 0.1 and dt < 0.001 * diameter, (ang, dt)
        worst_a, worst_t = max(worst_a, ang), max(worst_t, dt)
    assert np.abs(at_oracle_scores - OS).max(initial=0.0) <= 1e-5, (at_oracle_scores, OS)
    return worst_a, worst_t, float(np.abs(np.asarray(S) - OS).max(initial=0.0))


def oracle_pose_scores(rs, OP, n_resampled):
    """The GPU's dense score (inliers of the resampled model / N_re, clipped
    to 1, ppf.cpp:479-482) at each of the oracle's poses."""
    if len(OP) == 0:
        return np.zeros(0, np.float32)
    D4 = np.array([np.vstack([Q, [0, 0, 0, 1]]) for Q in OP], np.float32).reshape(-1, 4, 4)
    diameter = rs._det.model_info().diameter
    inl = rs.inliers(D4, dense=True, dist=np.float32(diameter) * np.float32(0.01))
    return np.minimum(inl.astype(np.float32) / np.float32(n_resampled), np.float32(1))


def check_votes(det, O, pts, nrm, n_keypoints, seed):
    """Sampled key points of a sampled scene: the GPU vote stage's
    accumulators and top-3 candidates against the oracle's vote loop and
    peak extraction on identical inputs.  Returns (voted, checked)."""
    info = det.model_info()
    sp, sn, _, _ = det.sampled_model()
    D = np.float32(info.diameter)
    ang = TWO_PI_F / np.float32(info.angleNum)
    dist = D * np.float32(0.04)
    ot = O.OracleTable(sp, sn, ang, dist)
    refNum = info.nSampled
    thr = np.float32(refNum) * np.float32(0.2)
    kps = np.sort(rng(seed).choice(len(pts), n_keypoints, replace=False)).astype(np.int32)
    cand, _, acc = det.vote(pts, nrm, kps, minScore=0.2, dump=len(kps))
    voted = 0
    for t, k in enumerate(kps):
        d = pts - pts[k]
        d2 = ((np.float32(0) + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        nb = np.nonzero(d2 < D * D)[0]
        if np.float32(len(nb)) < thr:
            assert (cand[t] == -1).all()
            continue
        nb = nb[nb != k]
        oacc = O.vote_keypoint(ot, refNum, info.angleNum, pts[k], nrm[k], pts[nb], nrm[nb], ang,
                               dist)
        assert np.array_equal(acc[t], oacc[:, 1:]), f"key point {k}"
        got = [tuple(c) for c in cand[t] if c[0] >= 0]
        assert got == [(a, b, int(v)) for a, b, v in O.peaks(oacc, refNum, info.angleNum, thr)]
        voted += 1
    return voted, len(kps)
