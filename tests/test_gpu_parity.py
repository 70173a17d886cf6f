"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer/byte/index results (hash keys, tables, neighbour sets, samples, vote
accumulators, peaks) must be bit-exact; floating-point results are checked
against the tolerances stated in each test.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

TWO_PI_F = np.float32(2.0) * np.float32(math.pi)


@pytest.fixture(scope="module")
def N():
    from surface_match_amd import _native
    _native.lib()
    return _native


def dev_math(N, op, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    b = None if b is None else np.ascontiguousarray(b, np.float32)
    out = np.zeros(len(a), np.uint32)
    N.check(N.lib().ppf_device_math(op, N.ptr(a), N.ptr(b), len(a), N.ptr(out)))
    return out


def test_l0_scalar_math_bitwise(N, oracle_mod):
    O = oracle_mod
    r = rng(1)
    a = np.concatenate([r.normal(size=4000) * 3, r.uniform(-1e4, 1e4, 1000),
                        [0.0, -0.0, 1.0, -1.0, 2.414213562373095, 0.41421356, 1e-30, -1e-30,
                         np.inf, -np.inf]]).astype(np.float32)
    b = np.concatenate([r.normal(size=4000) * 3, r.uniform(-1e4, 1e4, 1000),
                        [-1.0, -1.0, 0.0, -0.0, 1.0, 1.0, -1.0, -1.0, 1.0, 1.0]]).astype(np.float32)
    got = dev_math(N, 0, a)
    exp = np.array([O.lib().orc_atanf(float(x)) for x in a], np.float32).view(np.uint32)
    assert np.array_equal(got, exp)
    got = dev_math(N, 1, a, b)
    exp = np.array([O.lib().orc_atan2f(float(x), float(y)) for x, y in zip(a, b)],
                   np.float32).view(np.uint32)
    assert np.array_equal(got, exp)
    # IEEE division and square root are correctly rounded on gfx950
    assert np.array_equal(dev_math(N, 2, a, b), (a / b).view(np.uint32))
    aa = np.abs(a)
    assert np.array_equal(dev_math(N, 3, aa), np.sqrt(aa).view(np.uint32))
    # FP64 multiply + square root (the Jacobi eigen-solves' arithmetic)
    pa, pb = np.abs(a[:4000]), np.abs(b[:4000])
    d = np.sqrt(pa.astype(np.float64) * pb.astype(np.float64)).view(np.uint64)
    exp64 = ((d >> np.uint64(32)) ^ (d & np.uint64(0xFFFFFFFF))).astype(np.uint32)
    assert np.array_equal(dev_math(N, 7, pa, pb), exp64)
    # quantisation
    f = r.uniform(0, math.pi, 3000).astype(np.float32)
    st = np.full(3000, TWO_PI_F / np.float32(30), np.float32)
    assert np.array_equal(dev_math(N, 5, f, st), np.ceil(f / st).astype(np.int32).view(np.uint32))
    assert np.array_equal(dev_math(N, 6, f, st), (f / st).astype(np.int32).view(np.uint32))
    # murmur of (a bits, b bits, 0, 0)
    got = dev_math(N, 4, a, b)
    exp = np.array([O.lib().orc_hash_key4(np.array([x, y, 0, 0], np.uint32))
                    for x, y in zip(a.view(np.uint32), b.view(np.uint32))], np.uint32)
    assert np.array_equal(got, exp)


def _random_pairs(seed, n):
    r = rng(seed)
    p = r.uniform(-100, 100, (n, 3)).astype(np.float32)
    nn = r.normal(size=(n, 3))
    nn = (nn / np.linalg.norm(nn, axis=1, keepdims=True)).astype(np.float32)
    return p, nn


def test_l1_pair_features_bitwise(N, oracle_mod):
    O = oracle_mod
    p, n = _random_pairs(2, 3000)
    # adversarial partners: coincident point, parallel / antiparallel normals,
    # normals along +-x, coplanar pairs
    q1 = np.array([1.0, 2.0, 3.0], np.float32)
    m1 = np.array([0.0, 0.0, 1.0], np.float32)
    extra_p = np.array([q1, q1 + [5, 0, 0], q1 + [0, 5, 0], q1 + [0, 0, 7], q1 + [3, 4, 0]],
                       np.float32)
    extra_n = np.array([m1, m1, -m1, [1, 0, 0], [-1, 0, 0]], np.float32)
    pts = np.concatenate([p, extra_p]).astype(np.float32)
    nrm = np.concatenate([n, extra_n]).astype(np.float32)
    ang = TWO_PI_F / np.float32(30)
    dist = np.float32(5.89)
    for p1n1 in (np.concatenate([q1, m1]), np.array([0, 0, 0, 1, 0, 0], np.float32),
                 np.array([0, 0, 0, -1, 0, 0], np.float32), np.concatenate([p[0], n[0]])):
        p1n1 = p1n1.astype(np.float32)
        for tail in (0, 1):
            keys = np.zeros(len(pts), np.uint32)
            al = np.zeros(len(pts), np.float32)
            N.check(N.lib().ppf_pair_features(N.ptr(p1n1), N.ptr(pts), N.ptr(nrm), len(pts), ang,
                                              dist, tail, N.ptr(keys), N.ptr(al)))
            rt = O.frame(p1n1[:3], p1n1[3:])
            ek = np.array([O.ppf_key(p1n1[:3], p1n1[3:], pts[j], nrm[j], ang, dist, tail)[0]
                           for j in range(len(pts))], np.uint32)
            ea = np.array([O.alpha(rt, pts[j]) for j in range(len(pts))], np.float32)
            assert np.array_equal(keys, ek)
            assert np.array_equal(al.view(np.uint32), ea.view(np.uint32))


@pytest.fixture(scope="module")
def trained(N, workload1):
    from surface_match_amd import Detector, PointCloud
    mp, mn = workload1["model"]
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    return det


def test_l3_table_bitwise(trained, oracle_mod, workload1):
    O = oracle_mod
    sp, sn, rp, rn = trained.sampled_model()
    info = trained.model_info()
    keys, offs, refs, al = trained.table()
    ot = O.OracleTable(sp, sn, TWO_PI_F / np.float32(30), np.float32(info.diameter) * np.float32(0.04))
    assert np.array_equal(keys, ot.keys)
    assert np.array_equal(offs, ot.offs)
    assert np.array_equal(refs, ot.refs)
    assert np.array_equal(al.view(np.uint32), ot.alphas.view(np.uint32))
    # the sampled model is the oracle's training sample
    mp, mn = workload1["model"]
    om = O.train(mp, mn, 0.04)
    osp, osn = om.sampled()
    assert np.array_equal(osp, sp) and np.array_equal(osn, sn)
    orp, orn = om.resampled()
    assert np.array_equal(orp, rp) and np.array_equal(orn, rn)


def _vote_scene(workload1, oracle_mod, D):
    """A sampled scene with normals, as the vote stage sees it."""
    O = oracle_mod
    sc = workload1["scene"].points
    lo = sc.min(0)
    samp = O.voxel_sample(sc, np.float32(D) * np.float32(0.04), lo)
    n = np.full_like(sc, np.nan)
    valid = np.arange(len(sc), dtype=np.int32)
    pts, nrm = O.estimate_normals(sc, n, valid, samp, 10, 1, 0, (620.0, 100.0, 500.0))
    return pts[samp].copy(), nrm[samp].copy()


def test_l4_vote_accumulators_bitwise(trained, oracle_mod, workload1):
    O = oracle_mod
    info = trained.model_info()
    D = np.float32(info.diameter)
    pts, nrm = _vote_scene(workload1, O, D)
    sp, sn, _, _ = trained.sampled_model()
    ang = TWO_PI_F / np.float32(30)
    dist = D * np.float32(0.04)
    ot = O.OracleTable(sp, sn, ang, dist)
    r = rng(5)
    kps = np.sort(r.choice(len(pts), 48, replace=False)).astype(np.int32)
    refNum = info.nSampled
    thr = np.float32(refNum) * np.float32(0.2)
    D2 = D * D
    expect = []
    for k in kps:
        d = pts - pts[k]
        d2 = ((np.float32(0) + d[:, 0] * d[:, 0]) + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        nb = np.nonzero(d2 < D2)[0]
        if np.float32(len(nb)) < thr:
            expect.append(None)
            continue
        nb = nb[nb != k]
        oacc = O.vote_keypoint(ot, refNum, 30, pts[k], nrm[k], pts[nb], nrm[nb], ang, dist)
        expect.append((oacc, O.peaks(oacc, refNum, 30, thr)))
    # every launch shape (workgroup size, XCD mapping, grid cell) votes the same
    base = trained.get_tuning()
    try:
        # split and fused kernels; a hit capacity of 40 sends part of the key
        # points through the fused fallback
        for vb, remap, div, split, cap, tiles, order in (
                (1024, 0, 4, 1, 2048, 1, 0), (256, 1, 2, 0, 2048, 2, 1), (512, 0, 3, 1, 40, 3, 1),
                (1024, 1, 2, 0, 2048, 1, 0), (256, 0, 4, 1, 2048, 4, 1), (512, 1, 3, 1, 2048, 2, 1),
                (512, 1, 3, 1, 40, 1, 1)):
            trained.set_tuning(voteBlock=vb, voteXcdRemap=remap, voteCellDiv=div,
                               voteSplit=split, voteHitCap=cap, voteTiles=tiles, voteOrder=order)
            cand, poses, acc = trained.vote(pts, nrm, kps, minScore=0.2, dump=len(kps))
            for t, k in enumerate(kps):
                if expect[t] is None:
                    assert (cand[t] == -1).all()
                    continue
                oacc, pk = expect[t]
                assert np.array_equal(acc[t], oacc[:, 1:]), f"keypoint {k} tuning {vb, remap, div, split, cap, tiles, order}"
                assert np.array_equal(acc[t].sum(1), oacc[:, 0])
                got = [tuple(c) for c in cand[t] if c[0] >= 0]
                assert got == [(a, b, int(v)) for a, b, v in pk], f"keypoint {k}"
    finally:
        trained.set_tuning(**base)


def test_l4_vote_counts_past_16_bits(N, oracle_mod):
    """A flat model voted from a dense flat scene: the centre key point's
    busiest accumulator cells pass 65535 votes, which the 16-bit counters of
    the one-tile split path cannot hold; the overflow is detected and the key
    point voted again with 32-bit counters, so every tuning still equals the
    oracle's accumulators (for a corner key point as well)."""
    from surface_match_amd import Detector, PointCloud
    O = oracle_mod
    g = np.arange(0, 100.01, 2.0)
    X, Y = np.meshgrid(g, g, indexing="ij")
    mp = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1).astype(np.float32)
    mn = np.tile(np.array([0, 0, 1], np.float32), (len(mp), 1))
    det = Detector()
    det.trainModel(PointCloud(mp, mn), 0.04)
    sp, sn, _, _ = det.sampled_model()
    info = det.model_info()
    D = np.float32(info.diameter)
    ang = TWO_PI_F / np.float32(30)
    dist = D * np.float32(0.04)
    ot = O.OracleTable(sp, sn, ang, dist)
    r = rng(3)
    ns = 60000
    pts = np.concatenate([[[50, 50, 0], [2, 3, 0]],
                          np.stack([r.uniform(0, 100, ns), r.uniform(0, 100, ns),
                                    np.zeros(ns)], 1)]).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (len(pts), 1))
    kps = np.array([0, 1], np.int32)
    refNum = info.nSampled
    expect = []
    for k in kps:
        nb = np.array([j for j in range(len(pts)) if j != k], np.int64)
        expect.append(O.vote_keypoint(ot, refNum, 30, pts[k], nrm[k], pts[nb], nrm[nb], ang,
                                      dist))
    print("largest cells", [int(e[:, 1:].max()) for e in expect], flush=True)
    assert expect[0][:, 1:].max() > 65535
    base = det.get_tuning()
    try:
        for split, tiles, cap in ((1, 1, 1 << 17), (1, 2, 1 << 17), (0, 1, 2048), (1, 1, 2048)):
            det.set_tuning(voteSplit=split, voteTiles=tiles, voteHitCap=cap)
            cand, poses, acc = det.vote(pts, nrm, kps, minScore=0.2, dump=len(kps))
            for t in range(len(kps)):
                assert np.array_equal(acc[t], expect[t][:, 1:]), (t, split, tiles, cap)
    finally:
        det.set_tuning(**base)


def test_l2_voxel_and_search(N, oracle_mod, workload1):
    O = oracle_mod
    sc = workload1["scene"].points
    lo = sc.min(0).astype(np.float32)
    for step in (5.89, 1.47, 18.6):
        out = np.zeros(len(sc), np.int32)
        nout = C.c_size_t(0)
        N.check(N.lib().ppf_voxel_sample(N.ptr(sc), len(sc), None, 0, step, N.ptr(lo), N.ptr(out),
                                         C.byref(nout)))
        assert np.array_equal(out[:nout.value], O.voxel_sample(sc, step, lo))
    r = rng(7)
    q = sc[r.choice(len(sc), 300, replace=False)] + r.normal(scale=0.5, size=(300, 3)).astype(np.float32)
    q = q.astype(np.float32)
    r2 = np.float32(6.0) ** 2
    counts = np.zeros(len(q), np.int32)
    ids = np.zeros(len(q) * 400, np.int32)
    N.check(N.lib().ppf_radius_search(N.ptr(sc), len(sc), N.ptr(q), len(q), r2, 3.0,
                                      N.ptr(counts), N.ptr(ids), len(ids)))
    g = O.OracleGrid(sc, 4.0)
    off = 0
    for t in range(len(q)):
        exp = g.radius(q[t], r2)
        assert counts[t] == len(exp)
        assert np.array_equal(ids[off:off + counts[t]], exp)
        off += counts[t]
    k = 10
    kid = np.zeros(len(q) * k, np.int32)
    kd = np.zeros(len(q) * k, np.float32)
    N.check(N.lib().ppf_knn_search(N.ptr(sc), len(sc), N.ptr(q), len(q), k, 2.0, N.ptr(kid),
                                   N.ptr(kd)))
    for t in range(len(q)):
        eid, ed = g.knn(q[t], k)
        assert np.array_equal(kid[t * k:(t + 1) * k], eid)
        assert np.array_equal(kd[t * k:(t + 1) * k].view(np.uint32), ed.view(np.uint32))


def test_l5_normals(N, oracle_mod, workload1):
    O = oracle_mod
    sc = workload1["scene"].points
    lo = sc.min(0)
    samp = O.voxel_sample(sc, 5.89, lo)
    nrm0 = np.full_like(sc, np.nan)
    valid = np.arange(len(sc), dtype=np.int32)
    vp = np.array([620, 100, 500], np.float32)
    for smooth in (0, 1):
        ep, en = O.estimate_normals(sc, nrm0, valid, samp, 10, smooth, 0, vp)
        gp, gn = sc.copy(), nrm0.copy()
        N.check(N.lib().ppf_estimate_normals(N.ptr(gp), N.ptr(gn), len(sc), N.ptr(valid),
                                             len(valid), N.ptr(samp), len(samp), 10, smooth, 0,
                                             N.ptr(vp)))
        # positions and normals: bit-exact (same kNN order, same FP32 sums,
        # same FP64 Jacobi sequence on both sides)
        assert np.array_equal(np.isnan(gp), np.isnan(ep))
        m = ~np.isnan(ep[:, 0])
        assert np.array_equal(gp[m], ep[m])
        mn = ~np.isnan(en[:, 0])
        assert np.array_equal(mn, ~np.isnan(gn[:, 0]))
        assert np.array_equal(gn[mn], en[mn]), f"smooth={smooth}"


def test_match_workload1_poses(trained, oracle_mod, workload1):
    from surface_match_amd import PointCloud, MatchParam, synth
    O = oracle_mod
    sc = workload1["scene"]
    info = trained.model_info()
    D = info.diameter
    cloud = PointCloud(sc.points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    poses, scores = trained.matchScene(cloud, 0.04, 0.1, 0.2, MatchParam(numMatches=3))
    assert len(poses) == 3
    for P in poses:
        k, ang, dt = synth.pose_error(P, sc.gt_poses, D)
        assert ang < 0.5 and dt < 0.005 * D, (k, ang, dt)
    mp, mn = workload1["model"]
    om = O.train(mp, mn, 0.04)
    opos, osc, _ = O.match(om, sc.points, None, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT,
                           param=O.default_match_param(numMatches=3))
    # same instances found, poses within 0.1 deg / 0.1% D of the FP64 oracle refinement
    for P, s in zip(poses, scores):
        errs = [synth.pose_error(P, [np.vstack([Q, [0, 0, 0, 1]])], D) for Q in opos]
        best = min(errs, key=lambda e: e[1] + e[2])
        assert best[1] < 0.1 and best[2] < 0.001 * D, best
    print("gpu scores", scores, "oracle scores", osc, trained.last_stats)


@pytest.mark.parametrize("sparse,dense", [(1, 0), (0, 1), (0, 0), (1, 1)])
def test_match_refinement_switches(trained, oracle_mod, workload1, sparse, dense):
    """sparsePoseRefinement / densePoseRefinement on and off (ppf.cpp:423-495):
    without the sparse refinement the results are the clustered poses scored
    by their votes (equal to the oracle's bit for bit), the dense refinement
    only follows a sparse one; refined poses within 0.1 deg / 0.1 % D of the
    FP64 oracle refinement and scores within 1e-5.  The detector's helper
    stream prepares the refinement's inputs for every combination with
    refinement, so each case also runs that path."""
    from surface_match_amd import PointCloud, MatchParam, synth
    O = oracle_mod
    sc = workload1["scene"]
    D = trained.model_info().diameter
    cloud = PointCloud(sc.points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    poses, scores = trained.matchScene(
        cloud, 0.04, 0.1, 0.2,
        MatchParam(numMatches=3, sparsePoseRefinement=sparse, densePoseRefinement=dense))
    mp, mn = workload1["model"]
    om = O.train(mp, mn, 0.04)
    op = O.default_match_param(numMatches=3)
    op.sparsePoseRefinement, op.densePoseRefinement = sparse, dense
    opos, osc, _ = O.match(om, sc.points, None, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT, param=op,
                           threads=8)
    assert len(poses) == len(opos) == 3
    if sparse:
        assert np.abs(np.asarray(scores) - osc).max() <= 1e-5, (scores, osc)
        assert np.all(np.asarray(scores) <= 1.0)
    else:
        assert np.array_equal(np.asarray(scores, np.float32), osc), (scores, osc)
        assert scores[0] > 1.0  # vote counts, not inlier fractions
    for P in poses:
        errs = [synth.pose_error(P, [np.vstack([Q, [0, 0, 0, 1]])], D) for Q in opos]
        _, ang, dt = min(errs, key=lambda e: e[1] + e[2])
        assert ang < 0.1 and dt < 0.001 * D, (sparse, dense, ang, dt)


def test_match_degenerate_and_translated_scenes(trained, oracle_mod, workload1):
    """ppf.cpp:167-168: an empty scene is an error; scenes with no valid
    point, or too few to give a key point, end with no result (the oracle
    too).  The detector still matches afterwards: the scene moved by 1e4 mm
    along every axis gives the oracle's results (same float arithmetic at
    that magnitude), poses within 0.1 deg / 0.1 % D and scores within 1e-5."""
    from surface_match_amd import PointCloud, MatchParam, synth
    O = oracle_mod
    sc = workload1["scene"]
    D = trained.model_info().diameter
    mp, mn = workload1["model"]
    om = O.train(mp, mn, 0.04)
    prm = MatchParam(numMatches=3)
    with pytest.raises(RuntimeError, match="empty scene in matchScene"):
        trained.matchScene(PointCloud(np.zeros((0, 3), np.float32)), 0.04, 0.1, 0.2, prm)
    for pts in (np.full((1000, 3), np.nan, np.float32), sc.points[:5].copy(),
                sc.points[:1].copy()):
        poses, scores = trained.matchScene(PointCloud(pts), 0.04, 0.1, 0.2, prm)
        _, osc, _ = O.match(om, pts, None, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT,
                            param=O.default_match_param(numMatches=3), threads=8)
        assert len(poses) == len(scores) == len(osc) == 0
    moved = (sc.points + np.float32(1e4)).astype(np.float32)
    cloud = PointCloud(moved)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    poses, scores = trained.matchScene(cloud, 0.04, 0.1, 0.2, prm)
    opos, osc, _ = O.match(om, moved, None, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT,
                           param=O.default_match_param(numMatches=3), threads=8)
    assert len(poses) == len(opos) == 3
    assert np.abs(np.asarray(scores) - osc).max() <= 1e-5, (scores, osc)
    shift = np.eye(4)
    shift[:3, 3] = 1e4
    for P in poses:
        errs = [synth.pose_error(P, [np.vstack([Q, [0, 0, 0, 1]])], D) for Q in opos]
        _, ang, dt = min(errs, key=lambda e: e[1] + e[2])
        assert ang < 0.1 and dt < 0.001 * D, (ang, dt)
        # and they are poses of the scene's instances, moved
        _, ang, dt = synth.pose_error(P, [shift @ G for G in sc.gt_poses], D)
        assert ang < 1.0 and dt < 0.01 * D, (ang, dt)


def test_match_scene_with_normals(trained, oracle_mod, workload1):
    """A scene that brings its own normals (ppf.cpp:170, 208-211): points
    with a NaN normal are dropped, the rest normalised instead of estimated,
    and the dense refinement keeps them.  Normals: the oracle's kNN
    estimates scaled by 0.5..3, NaN on 1 % of the points.  Same results as
    the oracle: poses within 0.1 deg / 0.1 % D; the score gate is the GPU's
    scoring at the oracle's pose, equal to the oracle's score within 1e-5
    (at the GPU's own pose, a rounding-level different one, a point at the
    scoring radius may flip, so the own-pose scores are held to one
    inlier)."""
    from surface_match_amd import PointCloud, MatchParam, ResidentScene, synth
    O = oracle_mod
    sc = workload1["scene"]
    D = trained.model_info().diameter
    pts = sc.points
    valid = np.arange(len(pts), dtype=np.int32)
    _, nrm = O.estimate_normals(pts, np.full_like(pts, np.nan), valid, valid, 10, 0, 0,
                                synth.VIEWPOINT)
    r = rng(11)
    nrm = (nrm * r.uniform(0.5, 3.0, (len(nrm), 1))).astype(np.float32)
    nrm[r.choice(len(nrm), len(nrm) // 100, replace=False)] = np.nan
    cloud = PointCloud(pts, nrm)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    rs = ResidentScene(trained, cloud)
    poses, scores, _ = rs.match(0.04, 0.1, 0.2, MatchParam(numMatches=3))
    om = O.train(*workload1["model"], 0.04)
    opos, osc, _ = O.match(om, pts, nrm, 0.04, 0.1, 0.2, vp=synth.VIEWPOINT,
                           param=O.default_match_param(numMatches=3), threads=8)
    assert len(poses) == len(opos) == 3
    info = trained.model_info()
    OP4 = np.array([np.vstack([Q, [0, 0, 0, 1]]) for Q in opos], np.float32)
    inl = rs.inliers(OP4, dense=True, dist=np.float32(D) * np.float32(0.01))
    at_oracle = np.minimum(inl.astype(np.float32) / np.float32(info.nReSampled), np.float32(1))
    assert np.abs(at_oracle - osc).max() <= 1e-5, (at_oracle, osc)
    one = 1.0 / info.nReSampled
    assert np.abs(np.asarray(scores) - osc).max() <= one + 1e-6, (scores, osc)
    for P in poses:
        errs = [synth.pose_error(P, [np.vstack([Q, [0, 0, 0, 1]])], D) for Q in opos]
        _, ang, dt = min(errs, key=lambda e: e[1] + e[2])
        assert ang < 0.1 and dt < 0.001 * D, (ang, dt)
        _, ang, dt = synth.pose_error(P, sc.gt_poses, D)
        assert ang < 1.0 and dt < 0.01 * D, (ang, dt)


def test_save_load_roundtrip(trained, tmp_path):
    from surface_match_amd import Detector
    f = tmp_path / "m.model"
    trained.save(str(f))
    d2 = Detector()
    d2.load(str(f))
    a, b = trained.table(), d2.table()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    s1, s2 = trained.sampled_model(), d2.sampled_model()
    for x, y in zip(s1, s2):
        assert np.array_equal(x, y)
    # byte-identical re-save
    f2 = tmp_path / "m2.model"
    d2.save(str(f2))
    assert f.read_bytes() == f2.read_bytes()


def test_autotune_keeps_results(trained, workload1):
    """The autotuner only changes launch shapes: votes are identical and the
    refined poses agree to rounding."""
    from surface_match_amd import PointCloud, MatchParam, ResidentScene, synth
    sc = workload1["scene"]
    cloud = PointCloud(sc.points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    rs = ResidentScene(trained, cloud)
    base = trained.get_tuning()
    try:
        p0, s0, st0 = rs.match(0.04, 0.1, 0.2, MatchParam(numMatches=3))
        rep = rs.autotune(0.04, 0.1, 0.2, MatchParam(numMatches=3))
        assert rep["configsTried"] == (2 + 3 + 3 + 4 + 2 + 2) + 2 * 5 + 4
        assert trained.get_tuning() == rep["best"]
        p1, s1, st1 = rs.match(0.04, 0.1, 0.2, MatchParam(numMatches=3))
        for k in ("nPairs", "nKeyHits", "nVotes", "nPoses", "nHypotheses"):
            assert st0[k] == st1[k], k
        assert len(p0) == len(p1)
        D = trained.model_info().diameter
        for a, b in zip(p0, p1):
            assert synth.rotation_angle_deg(a[:3, :3].T @ b[:3, :3]) < 0.05
            assert np.linalg.norm(a[:3, 3] - b[:3, 3]) < 5e-4 * D
        assert np.allclose(s0, s1, atol=2.0 / trained.model_info().nReSampled)
        print("autotune", rep)
    finally:
        trained.set_tuning(**base)


def test_batched_scenes_equal_individual(trained):
    """Several scenes matched together (one vote launch, one batched
    refinement per stage) give every scene exactly its single-scene result,
    also with a scene of no valid point first and a one-point scene between
    (no key point: no result)."""
    from surface_match_amd import PointCloud, MatchParam, ResidentScene, synth
    scenes = synth.batch_scenes(0x5EED0003, n_scenes=4, lo=20000, hi=80000)
    clouds = [np.full((500, 3), np.nan, np.float32)] + [sc.points for sc in scenes]
    clouds.insert(3, scenes[0].points[:1].copy())
    rs = []
    for pts in clouds:
        c = PointCloud(pts)
        c.viewPoint = np.array(synth.VIEWPOINT, np.float32)
        rs.append(ResidentScene(trained, c))
    param = MatchParam(numMatches=3)
    many, st = ResidentScene.match_many(rs, 0.04, 0.1, 0.2, param)
    assert len(many) == len(rs)
    total = 0
    for r, (P, S) in zip(rs, many):
        P1, S1, st1 = r.match(0.04, 0.1, 0.2, param)
        assert np.array_equal(S, S1)
        assert np.array_equal(P, P1)
        total += st1["nKeypoints"]
    assert st["nKeypoints"] == total
    assert len(many[0][1]) == 0 and len(many[3][1]) == 0
    assert all(len(S) > 0 for i, (_, S) in enumerate(many) if i not in (0, 3))
    print("batched scenes", [len(S) for _, S in many], st["msTotal"])


def test_scene_from_ply_equals_scene_from_arrays(trained, workload1, tmp_path):
    """A scene read from a PLY file (ppf_scene_upload_ply) matches exactly as
    the same points handed over as arrays."""
    from surface_match_amd import PointCloud, MatchParam, ResidentScene, synth, writePLY
    pts = workload1["scene"].points
    path = tmp_path / "scene.ply"
    assert writePLY(str(path), PointCloud(pts))
    cloud = PointCloud(pts)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    param = MatchParam(numMatches=3)
    P1, S1, _ = ResidentScene(trained, cloud).match(0.04, 0.1, 0.2, param)
    P2, S2, _ = ResidentScene(trained, ply=str(path), viewPoint=synth.VIEWPOINT).match(
        0.04, 0.1, 0.2, param)
    assert np.array_equal(P1, P2) and np.array_equal(S1, S2)
