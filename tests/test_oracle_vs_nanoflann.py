"""The oracle's grid search and greedy sampling against the reference's own
kd-tree layer (nanoflann 1.5.5 as vendored and modified by the reference),
built from the reference headers into oracle/_ref by `make -C oracle ref`.
Skipped where that harness has not been built."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import rng

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libkdtree_ref.so")
pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def K():
    L = C.CDLL(REF)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.kdref_radius.argtypes = [fp, C.c_int, fp, C.c_float, ip, C.c_int]
    L.kdref_knn.argtypes = [fp, C.c_int, fp, C.c_int, ip, fp]
    L.kdref_greedy.argtypes = [fp, C.c_int, C.c_float, ip, ip]
    return L


def _cloud(seed, n=6000):
    from surface_match_amd import synth
    sc = synth.make_scene(seed, 2, n // 2, 0.3, 0.1, 60.0, plane_points=n // 4)
    return sc.points


def test_radius_sets_match_nanoflann(K, oracle_mod):
    pts = _cloud(60)
    g = oracle_mod.OracleGrid(pts, 3.0)
    r = rng(61)
    for q in pts[r.choice(len(pts), 200, replace=False)]:
        for r2 in (np.float32(4.0), np.float32(49.0), np.float32(400.0)):
            ids = np.zeros(len(pts), np.int32)
            n = K.kdref_radius(pts, len(pts), np.ascontiguousarray(q), r2, ids, len(ids))
            assert np.array_equal(np.sort(ids[:n]), g.radius(q, r2))


def test_knn_matches_nanoflann(K, oracle_mod):
    pts = _cloud(62)
    g = oracle_mod.OracleGrid(pts, 2.0)
    r = rng(63)
    ties = 0
    for q in pts[r.choice(len(pts), 300, replace=False)]:
        ids = np.zeros(10, np.int32)
        d2 = np.zeros(10, np.float32)
        K.kdref_knn(pts, len(pts), np.ascontiguousarray(q), 10, ids, d2)
        oid, od = g.knn(q, 10)
        assert np.array_equal(d2, od)          # same distances, bit for bit
        if len(np.unique(od)) == len(od):
            assert np.array_equal(ids, oid)    # same ids unless distances tie
        else:
            ties += 1
    assert ties < 5


def test_greedy_sampling_matches_reference_loop(K, oracle_mod):
    """samplePointCloud (util.cpp:107-139) single-threaded over the tree order."""
    pts = _cloud(64)
    order = np.zeros(len(pts), np.int32)
    out = np.zeros(len(pts), np.int32)
    n = K.kdref_greedy(pts, len(pts), np.float32(2.5), order, out)
    mine = np.zeros(len(pts), np.int32)
    m = oracle_mod.lib().orc_greedy_sample(len(pts), pts, order, len(order), np.float32(2.5), mine)
    assert np.array_equal(out[:n], mine[:m])
    # voxel-grid sampling (the GPU's replacement) keeps a comparable density
    vox = oracle_mod.voxel_sample(pts, np.float32(2.5), pts.min(0))
    assert 0.8 < len(vox) / n < 2.5
