"""The oracle's ICP::regist rules (icp.cpp:214-278) and estimateNormal's
smoothing rule (util.cpp:292-305), pinned by hand-made cases whose outcome
follows from the reference's code alone.  Each case is chosen so that a
plausible mistake in the restatement fails it:

ICP (icp.cpp:233-272; ConvergenceResult() = ITER, mse FLT_MAX, rate 1):
  * fewer than 6 correspondences -> NO_CORRESPONDS, nothing moves;
  * exactly 6 -> solved, but not converged (`validPairs > 6`);
  * 7 -> converged (so `>=` would fail the 6-pair case);
  * mse < mseMin stops with MSE after one iteration;
  * the first rate is (FLT_MAX - mse) / FLT_MAX == 1, which does not stop
    at tolerance 1 (`<`, not `<=`); the second rate does (CONVERGE_RATE);
  * `iterations++ >= iterations` never fires inside the loop: a run that
    never stops early ends with the ITER type after exactly `iterations`;
  * a near-identity initial pose (every entry within 1e-5 of I, Eigen's
    isIdentity) leaves the source untransformed while the result still
    composes it: with src == dst the result keeps the 9e-6 offset; at 2e-5
    the source moves and the refinement removes the offset.
Smoothing: the query point's own PCA normal plus its neighbours' normals,
each added when its dot with the query's normal is > 0 and subtracted
otherwise (self included, a perpendicular neighbour subtracted).
"""
import ctypes as C

import numpy as np
import pytest


def _cloud(k):
    from surface_match_amd import synth
    p, n = synth.make_model(4000, 77)
    # well-spread points: every 37th of the model
    idx = np.arange(0, 37 * k, 37)[:k]
    return np.ascontiguousarray(p[idx], np.float32), np.ascontiguousarray(n[idx], np.float32)


def _regist(O, src, srcn, dst, dstn, init, iters=5, reject=1e3, mse_min=0.0, mse_max=1e3,
            tol=0.05, mode=0):
    g = O.OracleGrid(dst, 5.0)
    crit = O.Criteria(iters, reject, mse_min, mse_max, tol)
    res = O.IcpResult()
    O.lib().orc_icp_regist(len(src), src, srcn, C.byref(g._g), np.ascontiguousarray(dstn),
                           np.ascontiguousarray(np.asarray(init, np.float32).reshape(12)),
                           C.byref(crit), mode, C.byref(res))
    return res, np.array(res.pose, np.float32).reshape(3, 4)


I34 = np.eye(4, dtype=np.float32)[:3]


@pytest.mark.parametrize("mode", [0, 1])
def test_icp_correspondence_count_rules(oracle_mod, mode):
    O = oracle_mod
    for k, want_type, want_conv in ((5, 3, 0), (6, 1, 0), (7, 1, 1)):
        p, n = _cloud(k)
        res, pose = _regist(O, p, n, p, n, I34, mse_min=1e-3, mode=mode)
        assert res.type == want_type, (k, res.type)
        assert res.converged == want_conv, (k, res.converged)
        assert res.iterations == (0 if k < 6 else 1), (k, res.iterations)
        if k < 6:  # nothing solved, the initial pose stays
            assert np.array_equal(pose, I34)


def test_icp_rate_and_iteration_rules(oracle_mod):
    O = oracle_mod
    from surface_match_amd import synth
    p, n = _cloud(60)
    r = np.random.Generator(np.random.PCG64(5))
    q = (p + r.normal(scale=0.05, size=p.shape)).astype(np.float32)
    R = synth.random_rotation(r)
    w = np.radians(2.0) * R[0]                      # a 2 degree turn about a random axis
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]) / th
    Rot = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    init = np.hstack([Rot, [[0.3], [-0.2], [0.1]]]).astype(np.float32)
    # tolerance 1: the first rate is exactly 1 (no stop), the second < 1
    res, _ = _regist(O, p, n, q, n, init, iters=10, tol=1.0)
    assert res.type == 2 and res.iterations == 2, (res.type, res.iterations)
    # tolerance 0, no mse floor, and a second step that still improves (rate
    # in (0, 1)): the loop ends at `iterations` with the ITER type
    res, _ = _regist(O, p, n, q, n, init, iters=2, tol=0.0)
    assert res.type == 0 and res.iterations == 2, (res.type, res.iterations)
    assert 0.0 < res.convergeRate < 1.0


def test_icp_near_identity_initial_pose(oracle_mod):
    O = oracle_mod
    p, n = _cloud(60)
    near = I34.copy()
    near[0, 3] = 9e-6                                 # within isIdentity's 1e-5
    res, pose = _regist(O, p, n, p, n, near, tol=0.0, iters=3)
    assert abs(pose[0, 3] - 9e-6) < 1e-6, pose[0, 3]  # composed, never applied
    far = I34.copy()
    far[0, 3] = 2e-5                                  # beyond it: the source moves
    res, pose = _regist(O, p, n, p, n, far, tol=0.0, iters=3)
    assert abs(pose[0, 3]) < 5e-6, pose[0, 3]


def _smoothing_case():
    """A query on the plane z = 0 with four neighbours at distance 1 whose
    normals are given (finite normals are not recomputed, util.cpp:239-241):
    the query's own PCA normal is +-z exactly."""
    pts = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0]], np.float32)
    nrm = np.array([[np.nan] * 3,
                    [0.6, 0.0, 0.8],      # +z side: added
                    [0.0, -0.6, -0.8],    # -z side: subtracted
                    [0.0, 0.0, 1.0],      # added
                    [1.0, 0.0, 0.0]],     # perpendicular (dot 0): subtracted
                   np.float32)
    return pts, nrm


def test_normal_smoothing_rule(oracle_mod):
    O = oracle_mod
    pts, nrm = _smoothing_case()
    P, N = O.estimate_normals(pts, nrm, np.arange(5), np.array([0]), k=5, smooth=1, invert=0,
                              vp=np.array([0, 0, 10], np.float32))
    # self (+z) + n1 - n2 + n3 - n4, oriented towards the view point
    want = np.array([0.6 - 1.0, 0.6, 0.8 + 0.8 + 1.0 + 1.0])
    want /= np.linalg.norm(want)
    assert np.allclose(N[0], want, atol=1e-6), (N[0], want)
    # the mistakes this rules out give other vectors
    for wrong in ([0.6 + 1.0, 0.6, 3.6],              # `>=` (perpendicular added)
                  [0.6 - 1.0, 0.6, 2.6],              # self excluded
                  [0.6 + 1.0, -0.6, 2.0]):            # no sign alignment
        w = np.array(wrong) / np.linalg.norm(wrong)
        assert not np.allclose(want, w, atol=1e-3)
    # the query moved to its neighbourhood's centroid (util.cpp:269)
    assert np.allclose(P[0], [0, 0, 0])


@pytest.mark.gpu
def test_gpu_normal_smoothing_rule(oracle_mod):
    """The GPU's estimate_normals on the same hand case."""
    from surface_match_amd import _native as N
    pts, nrm = _smoothing_case()
    P, Nn = pts.copy(), nrm.copy()
    vp = np.array([0, 0, 10], np.float32)
    valid = np.arange(5, dtype=np.int32)
    q = np.array([0], np.int32)
    N.check(N.lib().ppf_estimate_normals(N.ptr(P), N.ptr(Nn), 5, N.ptr(valid), 5, N.ptr(q), 1, 5,
                                         1, 0, N.ptr(vp)))
    want = np.array([0.6 - 1.0, 0.6, 3.6])
    want /= np.linalg.norm(want)
    assert np.allclose(Nn[0], want, atol=1e-6), (Nn[0], want)
