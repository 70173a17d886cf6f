"""Pins of the oracle's stages (L2-L6) against brute force, library routines
and closed forms.  CPU only; sizes keep each test to about a second."""
import math

import numpy as np
import pytest

from conftest import rng

F32_PI = np.float32(math.pi)
TWO_PI_F = np.float32(2.0) * F32_PI
ANGLE_STEP = TWO_PI_F / np.float32(30)


def _tiny_model(n=41, seed=20):
    from surface_match_amd import synth
    p, nn = synth.make_model(n, seed)
    return p, nn


def test_table_vs_bruteforce(oracle_mod):
    """ppf.cpp:126-143 single-thread: for i (finite n_i), j != i (finite
    alpha): bucket[key_ij] gets (i, alpha_ij) in (i, j) order; partners j >=
    n - n%8 use the scalar-tail key."""
    O = oracle_mod
    p, n = _tiny_model()
    n[5] = np.nan  # a NaN normal drops every pair (5, j)
    dist = np.float32(5.0)
    for tail in (1, 0):
        buckets = {}
        N = len(p)
        for i in range(N):
            if np.isnan(n[i]).any():
                continue
            rt = O.frame(p[i], n[i])
            for j in range(N):
                if j == i:
                    continue
                a = O.alpha(rt, p[j])
                if np.isnan(a):
                    continue
                k = O.ppf_key(p[i], n[i], p[j], n[j], ANGLE_STEP, dist,
                              int(tail and j >= N - N % 8))[0]
                buckets.setdefault(k, []).append((i, a))
        t = O.OracleTable(p, n, ANGLE_STEP, dist, tail)
        assert list(t.keys) == sorted(buckets)
        for b, k in enumerate(t.keys):
            ent = buckets[int(k)]
            s, e = t.offs[b], t.offs[b + 1]
            assert list(t.refs[s:e]) == [i for i, _ in ent]
            assert np.array_equal(t.alphas[s:e], np.array([a for _, a in ent], np.float32))
        assert 5 not in set(t.refs.tolist())


def _vote_brute(O, t, refNum, kp, kn, nbp, nbn, dist):
    acc = np.zeros((refNum, 31), np.int64)
    rt = O.frame(kp, kn)
    table = {int(k): (t.offs[b], t.offs[b + 1]) for b, k in enumerate(t.keys)}
    for j in range(len(nbp)):
        a_s = O.alpha(rt, nbp[j])
        k = O.ppf_key(kp, kn, nbp[j], nbn[j], ANGLE_STEP, dist, 0)[0]
        if k not in table or np.isnan(a_s):
            continue
        s, e = table[k]
        size = e - s
        for x in range(s, e):
            tail = (x - s) >= size - size % 8
            b = O.lib().orc_angle_bin(float(t.alphas[x]), float(a_s), 30, int(tail))
            acc[t.refs[x], 0] += 1
            acc[t.refs[x], b + 1] += 1
    return acc


def _peaks_brute(acc, vote_threshold):
    """std::multiset<Candidate, vote-greater> of size <= 3 + the 0.95 filter."""
    thre = np.float32(vote_threshold) / np.float32(2)
    ms = []
    for i in range(acc.shape[0]):
        el = acc[i]
        if np.float32(el[0]) < thre:
            continue
        b = el[1:]
        j = int(np.argmax(b))            # first max
        v = int(b[j])
        if np.float32(v) < thre:
            continue
        v += b[29] if j == 0 else b[j - 1]
        v += b[0] if j == 29 else b[j + 1]
        pos = len(ms)
        while pos > 0 and ms[pos - 1][0] < v:   # upper_bound insert
            pos -= 1
        ms.insert(pos, (float(v), i, j))
        if len(ms) > 3:
            ms.pop()
    if not ms:
        return []
    thr = np.float32(np.float64(np.float32(ms[0][0])) * 0.95)
    return [(i, j, v) for v, i, j in ms if not np.float32(v) < thr]


def test_vote_and_peaks_vs_bruteforce(oracle_mod):
    O = oracle_mod
    from surface_match_amd import synth
    p, n = synth.make_model(300, 21)
    mp, mn = synth.make_model(60, 22)
    dist = np.float32(8.0)
    t = O.OracleTable(mp, mn, ANGLE_STEP, dist)
    r = rng(23)
    for trial in range(4):
        k = int(r.integers(len(p)))
        nb = np.array([j for j in range(len(p)) if j != k])
        acc = O.vote_keypoint(t, len(mp), 30, p[k], n[k], p[nb], n[nb], ANGLE_STEP, dist)
        exp = _vote_brute(O, t, len(mp), p[k], n[k], p[nb], n[nb], dist)
        assert np.array_equal(acc, exp)
        for thr in (2.0, 12.0, 40.0):
            assert O.peaks(acc, len(mp), 30, thr) == _peaks_brute(exp, thr)


def test_peaks_ties_and_wrap(oracle_mod):
    O = oracle_mod
    acc = np.zeros((5, 31), np.int32)
    # four refs reach vote 12: ref 1 peaks at bin 29 and wraps to bin 0,
    # ref 2 peaks at bin 0 and wraps to bin 29; among equal votes the lower
    # ref ranks first and the 4th (ref 3) falls off the size-3 multiset;
    # ref 4 sits exactly at the threshold (count 3 = thre is kept)
    for r_, bins in {0: {5: 10, 4: 2}, 1: {29: 10, 0: 2}, 2: {0: 9, 29: 2, 1: 1},
                     3: {7: 10, 8: 2}, 4: {2: 3}}.items():
        for b, v in bins.items():
            acc[r_, b + 1] = v
        acc[r_, 0] = sum(bins.values())
    got = O.peaks(acc, 5, 30, 6.0)
    assert got == [(0, 5, 12.0), (1, 29, 12.0), (2, 0, 12.0)]
    assert got == _peaks_brute(acc, 6.0)
    # the 0.95 filter drops candidates below 0.95 x best
    acc[0, 6] += 1
    acc[0, 0] += 1
    assert O.peaks(acc, 5, 30, 6.0) == [(0, 5, 13.0)] == _peaks_brute(acc, 6.0)


def test_grid_search_vs_bruteforce(oracle_mod):
    O = oracle_mod
    r = rng(30)
    pts = np.concatenate([r.uniform(-50, 50, (3000, 3)),
                          r.normal(scale=2, size=(500, 3))]).astype(np.float32)
    pts[7] = pts[8]  # exact duplicate: ties broken by lower id
    for cell in (1.0, 3.0, 17.0):
        g = O.OracleGrid(pts, cell)
        for q in np.concatenate([pts[:20], r.uniform(-60, 60, (20, 3))]).astype(np.float32):
            d = pts - q
            d2 = ((np.float32(0) + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            for r2 in (np.float32(4.0), np.float32(25.0)):
                assert np.array_equal(g.radius(q, r2), np.nonzero(d2 < r2)[0])
            ids, dd = g.knn(q, 10)
            order = np.lexsort((np.arange(len(pts)), d2))[:10]
            assert np.array_equal(ids, order) and np.array_equal(dd, d2[order])


def test_voxel_sample_vs_numpy(oracle_mod):
    O = oracle_mod
    r = rng(31)
    pts = r.uniform(-20, 20, (20000, 3)).astype(np.float32)
    lo = pts.min(0)
    for step in (np.float32(1.7), np.float32(4.0)):
        got = O.voxel_sample(pts, step, lo)
        v = np.floor((pts - lo) / step).astype(np.int64)
        c = lo + (v.astype(np.float32) + np.float32(0.5)) * step
        d = pts - c
        d2 = ((np.float32(0) + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        key = (v[:, 0] + (1 << 20)) << 42 | (v[:, 1] + (1 << 20)) << 21 | (v[:, 2] + (1 << 20))
        order = np.lexsort((np.arange(len(pts)), d2, key))
        first = order[np.r_[True, key[order][1:] != key[order][:-1]]]
        assert np.array_equal(got, np.sort(first))
        # one point per voxel, every voxel covered
        assert len(got) == len(np.unique(key))


def test_greedy_sample_is_poisson(oracle_mod):
    """util.cpp:107-139: kept points are >= step apart and every point is
    within step of a kept point (maximal), in the given order."""
    O = oracle_mod
    r = rng(32)
    pts = r.uniform(0, 30, (3000, 3)).astype(np.float32)
    order = r.permutation(len(pts)).astype(np.int32)
    out = np.zeros(len(pts), np.int32)
    n = O.lib().orc_greedy_sample(len(pts), pts, order, len(order), np.float32(3.0), out)
    kept = out[:n]
    d2 = ((pts[kept][:, None, :] - pts[kept][None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, 1e9)
    assert d2.min() >= 9.0 - 1e-3
    dall = ((pts[:, None, :] - pts[kept][None, :, :]) ** 2).sum(-1).min(1)
    assert dall.max() < 9.0 + 1e-3
    # the first point of the order is always kept
    assert kept[0] == order[0]


def test_normals_vs_eigh(oracle_mod):
    O = oracle_mod
    r = rng(33)
    # a noisy tilted plane and a sphere
    u = r.uniform(-30, 30, (4000, 2))
    nrm_true = np.array([0.3, -0.2, 0.93])
    nrm_true /= np.linalg.norm(nrm_true)
    a = np.cross(nrm_true, [1, 0, 0])
    a /= np.linalg.norm(a)
    b = np.cross(nrm_true, a)
    plane = u[:, :1] * a + u[:, 1:] * b + r.normal(scale=0.01, size=(4000, 1)) * nrm_true
    th = r.normal(size=(3000, 3))
    sphere = 20 * th / np.linalg.norm(th, axis=1, keepdims=True) + [100, 0, 0]
    pts = np.concatenate([plane, sphere]).astype(np.float32)
    n0 = np.full_like(pts, np.nan)
    valid = np.arange(len(pts), dtype=np.int32)
    q = np.arange(0, len(pts), 7, dtype=np.int32)
    vp = np.array([0, 0, 100], np.float32)
    P, Nn = O.estimate_normals(pts, n0, valid, q, 10, 0, 0, vp)
    g = O.OracleGrid(pts, 2.0)
    for i in q[:300]:
        ids, _ = g.knn(pts[i], 10)
        X = pts[ids].astype(np.float64)
        c = X.mean(0)
        w, V = np.linalg.eigh((X - c).T @ (X - c))
        e = V[:, np.argmin(np.abs(w))]
        assert abs(abs(np.dot(e, Nn[i])) - 1) < 1e-5
        assert np.allclose(P[i], c, atol=1e-4)   # point moved to the centroid
        # oriented toward the view point (util.cpp:313-322)
        assert np.dot(Nn[i], vp - P[i]) >= 0
    # plane normals match the plane
    pl = q[q < 4000]
    assert np.degrees(np.arccos(np.clip(np.abs(Nn[pl] @ nrm_true), -1, 1))).max() < 2.0
    # smoothing keeps unit length and does not move non-touched points
    P2, N2 = O.estimate_normals(pts, n0, valid, q, 10, 1, 0, vp)
    assert np.allclose(np.linalg.norm(N2[q], axis=1), 1, atol=1e-5)


def _pose_t(O, R, t, votes):
    from oracle import PoseT
    p = PoseT()
    P = np.zeros(12, np.float32)
    P.reshape(3, 4)[:, :3] = R
    P.reshape(3, 4)[:, 3] = t
    p.pose[:] = P.tolist()
    p.votes = votes
    q = np.zeros(4, np.float32)
    O.lib().orc_quat_from_rot(np.ascontiguousarray(R.astype(np.float32).ravel()), q)
    p.q[:] = q.tolist()
    p.angle = O.lib().orc_rot_angle(np.ascontiguousarray(R.astype(np.float32).ravel()))
    return p


def _cluster(O, poses, dthr, athr):
    import ctypes as C
    from oracle import PoseT
    arr = (PoseT * len(poses))(*poses)
    out = (PoseT * len(poses))()
    cl = np.zeros(len(poses), np.int32)
    nc = O.lib().orc_cluster_poses(arr, len(poses), dthr, athr, cl, out)
    return cl, [out[i] for i in range(nc)]


def test_clustering_hand_cases(oracle_mod):
    O = oracle_mod
    from surface_match_amd.synth import random_rotation
    r = rng(40)
    R = random_rotation(r).astype(np.float32)
    # identical poses -> one cluster, averaged pose equals the input
    ps = [_pose_t(O, R, [1, 2, 3], 5.0) for _ in range(4)]
    cl, cen = _cluster(O, ps, 1.0, 0.2)
    assert list(cl) == [0, 0, 0, 0] and len(cen) == 1
    P = np.array(cen[0].pose, np.float32).reshape(3, 4)
    assert np.allclose(P[:, :3], R, atol=1e-5) and np.allclose(P[:, 3], [1, 2, 3])
    assert cen[0].votes == 20.0
    # q and -q describe one rotation: the Markley mean is sign-blind
    p1 = _pose_t(O, R, [0, 0, 0], 3.0)
    p2 = _pose_t(O, R, [0, 0, 0], 2.0)
    p2.q[:] = [-v for v in p2.q]
    cl, cen = _cluster(O, [p1, p2], 1.0, 0.2)
    assert np.allclose(np.array(cen[0].pose).reshape(3, 4)[:, :3], R, atol=1e-5)
    # two separated groups -> two clusters, ordered by votes
    a = [_pose_t(O, R, [0, 0, 0], 1.0), _pose_t(O, R, [50, 0, 0], 9.0),
         _pose_t(O, R, [0.5, 0, 0], 2.0)]
    cl, cen = _cluster(O, a, 1.0, 0.2)
    assert list(cl) == [1, 0, 1] and [c.votes for c in cen] == [9.0, 3.0]
    # greedy order: a pose joins the FIRST matching centre (by votes), not the nearest
    b = [_pose_t(O, R, [0, 0, 0], 9.0), _pose_t(O, R, [0.9, 0, 0], 8.0),
         _pose_t(O, R, [0.6, 0, 0], 1.0)]
    cl, cen = _cluster(O, b, 1.0, 0.2)
    assert list(cl) == [0, 0, 0]


def test_nms_hand_cases(oracle_mod):
    O = oracle_mod
    from oracle import PoseT
    from surface_match_amd.synth import random_rotation
    R = random_rotation(rng(41)).astype(np.float32)
    ps = [_pose_t(O, R, [0, 0, 0], 5.0), _pose_t(O, R, [1, 0, 0], 9.0),
          _pose_t(O, R, [30, 0, 0], 3.0), _pose_t(O, R, [1.5, 0, 0], 1.0)]
    arr = (PoseT * 4)(*ps)
    out = (PoseT * 4)()
    n = O.lib().orc_nms_poses(arr, 4, np.zeros(3, np.float32), 2.0, out)
    # by votes: 9 (x=1) absorbs 5 (x=0) and 1 (x=1.5); 3 (x=30) stays
    assert n == 2 and out[0].votes == 15.0 and out[1].votes == 3.0


def test_icp_recovers_small_motion(oracle_mod):
    O = oracle_mod
    from oracle import Criteria, IcpResult
    import ctypes as C
    from surface_match_amd import synth
    p, n = synth.make_model(3000, 50)
    r = rng(51)
    ang = np.radians(2.0)
    ax = r.normal(size=3)
    ax /= np.linalg.norm(ax)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
    t = np.array([1.0, -0.5, 0.8])
    dst = (p @ R.T + t).astype(np.float32)
    dn = (n @ R.T).astype(np.float32)
    g = O.OracleGrid(dst, 3.0)
    crit = Criteria(30, 20.0, 1e-4, 5.0, 0.0)
    init = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    for mode in (0, 1):
        res = IcpResult()
        O.lib().orc_icp_regist(len(p), p, n, C.byref(g._g), dn, init, C.byref(crit), mode,
                               C.byref(res))
        P = np.array(res.pose, np.float64).reshape(3, 4)
        err = synth.rotation_angle_deg(P[:, :3].T @ R)
        assert err < 0.01 and np.linalg.norm(P[:, 3] - t) < 0.01, (mode, err, P[:, 3])
        assert res.converged


def test_match_end_to_end_small(oracle_mod):
    """trainModel + matchScene on a reduced workload recover the ground truth."""
    O = oracle_mod
    from surface_match_amd import synth
    w = synth.workload(1, scale=0.5)
    mp, mn = w["model"]
    m = O.train(mp, mn, 0.04)
    sc = w["scene"]
    poses, scores, st = O.match(m, sc.points, None, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT,
                                param=O.default_match_param(numMatches=3), threads=4)
    D = float(np.linalg.norm(np.array(m.boxMax) - np.array(m.boxMin)))
    assert len(poses) >= 2
    found = set()
    for P in poses:
        k, ang, dt = synth.pose_error(P, sc.gt_poses, D)
        assert ang < 1.0 and dt < 0.01 * D
        found.add(k)
    assert len(found) == len(poses)
    assert (np.diff(scores) <= 0).all()


def test_remove_nan_and_bounding_box_hand_cases(oracle_mod):
    """removeNan (util.cpp:86-105) and computeBoundingBox (util.cpp:178-207)
    on hand-made clouds: non-finite coordinates drop a point, non-finite
    normals drop it only when the cloud has normals; the box starts at the
    first valid point and a coordinate replaces a bound only when strictly
    beyond it, so among equal extremes (+0 and -0 included) the earliest
    point's bits stay."""
    O = oracle_mod
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    p = np.array([[1, 2, 3], [nan, 0, 0], [0, inf, 0], [-1, 5, 2], [4, -7, 9]], np.float32)
    n = np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1], [nan, 0, 1], [1, 0, 0]], np.float32)
    assert O.remove_nan(p).tolist() == [0, 3, 4]
    assert O.remove_nan(p, n).tolist() == [0, 4]
    assert O.bounding_box(p, [0, 3, 4]).tolist() == [-1, -7, 2, 4, 5, 9]
    # ties of +0 and -0: the earliest holder of the extreme keeps its sign
    z = np.float32(0.0)
    mz = np.float32(-0.0)
    q = np.array([[z, 1, 1], [mz, 2, 2], [3, 3, 3]], np.float32)
    b = O.bounding_box(q, [0, 1, 2])
    assert b[0] == 0 and not np.signbit(b[0])          # +0 first, -0 later: +0
    b = O.bounding_box(q, [1, 0, 2])
    assert b[0] == 0 and np.signbit(b[0])              # -0 first: -0
    q = np.array([[-1, -1, -1], [mz, 0, 0], [z, 0, 0]], np.float32)
    b = O.bounding_box(q, [0, 1, 2])
    assert b[3] == 0 and np.signbit(b[3])              # max: -0 reached first
    # a repeated extreme: the first index (visible through -0/+0 only, so
    # check that later equal values do not move the bound)
    q = np.array([[5, 0, 0], [mz, 0, 0], [z, 0, 0], [mz, 0, 0]], np.float32)
    b = O.bounding_box(q, [0, 1, 2, 3])
    assert b[0] == 0 and np.signbit(b[0]) and b[3] == 5
