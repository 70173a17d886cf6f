"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libppf_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libkdtree_ref.so")

fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class Table(C.Structure):
    _fields_ = [("nKeys", C.c_int), ("nEntries", C.c_int), ("keys", C.POINTER(C.c_uint32)),
                ("offs", C.POINTER(C.c_int)), ("refs", C.POINTER(C.c_uint32)),
                ("alphas", C.POINTER(C.c_float))]

    def arrays(self):
        k, e = self.nKeys, self.nEntries
        return (np.ctypeslib.as_array(self.keys, (max(k, 1),))[:k].copy(),
                np.ctypeslib.as_array(self.offs, (k + 1,)).copy(),
                np.ctypeslib.as_array(self.refs, (max(e, 1),))[:e].copy(),
                np.ctypeslib.as_array(self.alphas, (max(e, 1),))[:e].copy())


class Candidate(C.Structure):
    _fields_ = [("refId", C.c_int), ("angleId", C.c_int), ("vote", C.c_float)]


class Grid(C.Structure):
    _fields_ = [("n", C.c_int), ("pts", C.c_void_p), ("org", C.c_float * 3), ("cell", C.c_float),
                ("dim", C.c_int * 3), ("start", C.c_void_p), ("idx", C.c_void_p)]


class PoseT(C.Structure):
    _fields_ = [("pose", C.c_float * 12), ("votes", C.c_float), ("angle", C.c_float),
                ("q", C.c_float * 4)]


class Criteria(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rejectDist", C.c_float), ("mseMin", C.c_float),
                ("mseMax", C.c_float), ("tolerance", C.c_float)]


class IcpResult(C.Structure):
    _fields_ = [("pose", C.c_float * 12), ("type", C.c_int), ("mse", C.c_float),
                ("convergeRate", C.c_float), ("iterations", C.c_int), ("converged", C.c_int)]


class TrainParam(C.Structure):
    _fields_ = [("featDistanceStepRel", C.c_float), ("featAngleResolution", C.c_int),
                ("poseRefRelSamplingDistance", C.c_float), ("knnNormal", C.c_int),
                ("smoothNormal", C.c_int)]


class MatchParam(C.Structure):
    _fields_ = [("numMatches", C.c_int), ("knnNormal", C.c_int), ("smoothNormal", C.c_int),
                ("invertNormal", C.c_int), ("maxOverlapDistRel", C.c_float),
                ("maxOverlapDistAbs", C.c_float), ("sparsePoseRefinement", C.c_int),
                ("densePoseRefinement", C.c_int), ("poseRefNumSteps", C.c_int),
                ("poseRefDistThresholdRel", C.c_float), ("poseRefDistThresholdAbs", C.c_float),
                ("poseRefScoringDistRel", C.c_float), ("poseRefScoringDistAbs", C.c_float)]


def default_train_param():
    return TrainParam(0.04, 30, 0.01, 10, 1)


def default_match_param(**kw):
    p = MatchParam(1, 10, 1, 0, 0.5, 0.0, 1, 1, 5, 0.1, 0.0, 0.01, 0.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Model(C.Structure):
    _fields_ = [("samplingDistanceRel", C.c_float), ("param", TrainParam), ("nSampled", C.c_int),
                ("nReSampled", C.c_int), ("sampledP", C.POINTER(C.c_float)),
                ("sampledN", C.POINTER(C.c_float)), ("reSampledP", C.POINTER(C.c_float)),
                ("reSampledN", C.POINTER(C.c_float)), ("boxMin", C.c_float * 3),
                ("boxMax", C.c_float * 3), ("viewPoint", C.c_float * 3), ("table", Table)]

    def sampled(self):
        n = self.nSampled
        return (np.ctypeslib.as_array(self.sampledP, (n, 3)).copy(),
                np.ctypeslib.as_array(self.sampledN, (n, 3)).copy())

    def resampled(self):
        n = self.nReSampled
        return (np.ctypeslib.as_array(self.reSampledP, (n, 3)).copy(),
                np.ctypeslib.as_array(self.reSampledN, (n, 3)).copy())


class MatchStats(C.Structure):
    _fields_ = [("nSampled", C.c_int), ("nKeypoints", C.c_int), ("nPoses", C.c_int),
                ("nClusters", C.c_int), ("nNms", C.c_int), ("nVotedKeypoints", C.c_int),
                ("nPairs", C.c_long), ("nKeyHits", C.c_long), ("nVotes", C.c_long),
                ("timeMs", C.c_float * 8)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_atanf.restype = C.c_float
        L.orc_atanf.argtypes = [C.c_float]
        L.orc_atan2f.restype = C.c_float
        L.orc_atan2f.argtypes = [C.c_float, C.c_float]
        L.orc_murmur3_x86_32.restype = C.c_uint32
        L.orc_murmur3_x86_32.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32]
        L.orc_hash_key4.restype = C.c_uint32
        L.orc_hash_key4.argtypes = [up]
        L.orc_ppf_key.restype = C.c_uint32
        L.orc_ppf_key.argtypes = [fp, fp, fp, fp, C.c_float, C.c_float, C.c_int, fp, up]
        L.orc_frame.argtypes = [fp, fp, fp]
        L.orc_alpha.restype = C.c_float
        L.orc_alpha.argtypes = [fp, fp]
        L.orc_angle_bin.restype = C.c_int
        L.orc_angle_bin.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int]
        L.orc_build_table.argtypes = [C.c_int, fp, fp, C.c_float, C.c_float, C.c_int,
                                      C.POINTER(Table)]
        L.orc_free_table.argtypes = [C.POINTER(Table)]
        L.orc_vote_keypoint.argtypes = [C.POINTER(Table), C.c_int, C.c_int, fp, fp, C.c_int, fp,
                                        fp, C.c_float, C.c_float, ip]
        L.orc_peaks.restype = C.c_int
        L.orc_peaks.argtypes = [ip, C.c_int, C.c_int, C.c_float, C.POINTER(Candidate)]
        L.orc_pose.argtypes = [fp, fp, C.c_int, C.c_int, fp]
        L.orc_grid_build.argtypes = [C.POINTER(Grid), C.c_int, fp, C.c_void_p, C.c_int,
                                     C.c_float]
        L.orc_grid_free.argtypes = [C.POINTER(Grid)]
        L.orc_grid_radius.restype = C.c_int
        L.orc_grid_radius.argtypes = [C.POINTER(Grid), fp, C.c_float, ip, C.c_int]
        L.orc_grid_knn.restype = C.c_int
        L.orc_grid_knn.argtypes = [C.POINTER(Grid), fp, C.c_int, ip, fp]
        L.orc_voxel_sample.restype = C.c_int
        L.orc_voxel_sample.argtypes = [C.c_int, fp, C.c_void_p, C.c_int, C.c_float, fp, ip]
        L.orc_greedy_sample.restype = C.c_int
        L.orc_greedy_sample.argtypes = [C.c_int, fp, ip, C.c_int, C.c_float, ip]
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_unit_orthogonal.argtypes = [fp, fp]
        L.orc_estimate_normals_mls.restype = C.c_long
        L.orc_estimate_normals_mls.argtypes = [C.c_int, fp, fp, ip, C.c_int, ip, C.c_int,
                                               C.c_float, C.c_int, C.c_int, C.c_void_p]
        L.orc_philox4x32_10.argtypes = [up, up]
        L.orc_sample_mesh.restype = C.c_int
        L.orc_sample_mesh.argtypes = [C.c_int, fp, C.c_int, C.c_void_p, C.c_float, C.c_uint64,
                                      C.c_void_p, C.c_void_p, C.c_int]
        L.orc_estimate_normals.argtypes = [C.c_int, fp, fp, ip, C.c_int, ip, C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_void_p]
        L.orc_jacobi_eig.restype = C.c_int
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_jacobi_eig.argtypes = [C.c_int, dp, dp, dp]
        L.orc_quat_from_rot.argtypes = [fp, fp]
        L.orc_rot_from_quat.argtypes = [fp, fp]
        L.orc_rot_angle.restype = C.c_float
        L.orc_rot_angle.argtypes = [fp]
        L.orc_cluster_poses.restype = C.c_int
        L.orc_cluster_poses.argtypes = [C.POINTER(PoseT), C.c_int, C.c_float, C.c_float, ip,
                                        C.POINTER(PoseT)]
        L.orc_nms_poses.restype = C.c_int
        L.orc_nms_poses.argtypes = [C.POINTER(PoseT), C.c_int, fp, C.c_float, C.POINTER(PoseT)]
        L.orc_icp_regist.argtypes = [C.c_int, fp, fp, C.POINTER(Grid), fp, fp,
                                     C.POINTER(Criteria), C.c_int, C.POINTER(IcpResult)]
        L.orc_inliers.restype = C.c_int
        L.orc_inliers.argtypes = [C.c_int, fp, fp, C.POINTER(Grid), C.c_float]
        L.orc_p2pl_system.argtypes = [C.c_int, fp, dp, ip, fp, fp, dp, dp, C.c_int, dp]
        L.orc_train.argtypes = [C.c_int, fp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                C.POINTER(TrainParam), C.c_int, C.POINTER(Model)]
        L.orc_free_model.argtypes = [C.POINTER(Model)]
        L.orc_remove_nan.restype = C.c_int
        L.orc_remove_nan.argtypes = [C.c_int, fp, C.c_void_p, ip]
        L.orc_bounding_box.argtypes = [C.c_int, fp, ip, C.c_int, fp]
        L.orc_match.restype = C.c_int
        L.orc_match.argtypes = [C.POINTER(Model), C.c_int, fp, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_float, C.c_float, C.c_float,
                                C.POINTER(MatchParam), C.c_int, fp, fp, C.POINTER(MatchStats),
                                C.c_int]
        _lib = L
    return _lib


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- helpers

def ppf_key(p1, n1, p2, n2, angle_step, dist_step, tail=0):
    f = np.zeros(4, np.float32)
    q = np.zeros(4, np.uint32)
    k = lib().orc_ppf_key(f32(p1), f32(n1), f32(p2), f32(n2), angle_step, dist_step, tail, f, q)
    return int(k), f, q


def frame(p, n):
    rt = np.zeros(12, np.float32)
    lib().orc_frame(f32(p), f32(n), rt)
    return rt


def alpha(rt, p2):
    return float(lib().orc_alpha(f32(rt), f32(p2)))


class OracleTable:
    def __init__(self, pts, nrm, angle_step, dist_step, ref_tail=1):
        self._t = Table()
        pts, nrm = f32(pts), f32(nrm)
        lib().orc_build_table(len(pts), pts, nrm, angle_step, dist_step, ref_tail,
                              C.byref(self._t))
        self.keys, self.offs, self.refs, self.alphas = self._t.arrays()

    def __del__(self):
        try:
            lib().orc_free_table(C.byref(self._t))
        except Exception:
            pass


def vote_keypoint(table: OracleTable, ref_num, angle_num, kp, kn, nb_pts, nb_nrm, angle_step,
                  dist_step):
    acc = np.zeros(ref_num * (angle_num + 1), np.int32)
    nb_pts = f32(nb_pts).reshape(-1, 3)
    nb_nrm = f32(nb_nrm).reshape(-1, 3)
    lib().orc_vote_keypoint(C.byref(table._t), ref_num, angle_num, f32(kp), f32(kn),
                            len(nb_pts), nb_pts, nb_nrm, angle_step, dist_step, acc)
    return acc.reshape(ref_num, angle_num + 1)


def peaks(acc, ref_num, angle_num, vote_threshold):
    out = (Candidate * 3)()
    n = lib().orc_peaks(np.ascontiguousarray(acc, np.int32).ravel(), ref_num, angle_num,
                        vote_threshold, out)
    return [(out[i].refId, out[i].angleId, out[i].vote) for i in range(n)]


def pose(scene_rt, model_rt, angle_id, angle_num):
    o = np.zeros(12, np.float32)
    lib().orc_pose(f32(scene_rt), f32(model_rt), angle_id, angle_num, o)
    return o


class OracleGrid:
    def __init__(self, pts, cell, subset=None):
        self.pts = f32(pts)
        self.sub = None if subset is None else np.ascontiguousarray(subset, np.int32)
        self._g = Grid()
        lib().orc_grid_build(C.byref(self._g), len(self.pts), self.pts, _ptr(self.sub),
                             0 if self.sub is None else len(self.sub), cell)

    def radius(self, q, r2, cap=None):
        cap = cap or len(self.pts)
        out = np.zeros(cap, np.int32)
        n = lib().orc_grid_radius(C.byref(self._g), f32(q), r2, out, cap)
        return out[:min(n, cap)]

    def knn(self, q, k):
        ids = np.zeros(k, np.int32)
        d = np.zeros(k, np.float32)
        n = lib().orc_grid_knn(C.byref(self._g), f32(q), k, ids, d)
        return ids[:n], d[:n]

    def __del__(self):
        try:
            lib().orc_grid_free(C.byref(self._g))
        except Exception:
            pass


def voxel_sample(pts, step, box_min, subset=None):
    pts = f32(pts)
    sub = None if subset is None else np.ascontiguousarray(subset, np.int32)
    out = np.zeros(len(pts) if sub is None else len(sub), np.int32)
    n = lib().orc_voxel_sample(len(pts), pts, _ptr(sub), 0 if sub is None else len(sub), step,
                               f32(box_min), out)
    return out[:n]


def estimate_normals_mls(pts, nrm, valid, query, radius, order=2, invert=0, vp=None):
    """estimateNormalMLS (util.cpp:336-457), snapshot semantics:
    (points, normals, total neighbourhood size)."""
    pts = f32(pts).copy()
    nrm = f32(nrm).copy()
    vpa = None if vp is None else f32(vp)
    tot = lib().orc_estimate_normals_mls(len(pts), pts, nrm, np.ascontiguousarray(valid, np.int32),
                                         len(valid), np.ascontiguousarray(query, np.int32),
                                         len(query), radius, order, invert, _ptr(vpa))
    if tot < 0:
        raise ValueError("orc_estimate_normals_mls: order or radius out of range")
    return pts, nrm, tot


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32).copy()
    lib().orc_philox4x32_10(c, np.ascontiguousarray(key, np.uint32))
    return c


def sample_mesh(pts, faces, radius, seed):
    """sampleMesh (util.cpp:20-84) with Philox uniforms: (points, normals)."""
    pts = f32(pts).reshape(-1, 3)
    fc = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    L = lib()
    n = L.orc_sample_mesh(len(pts), pts, len(fc), _ptr(fc), radius, seed, None, None, 0)
    if n < 0:
        raise ValueError("orc_sample_mesh failed (%d)" % n)
    P = np.zeros((n, 3), np.float32)
    N = np.zeros((n, 3), np.float32)
    if n:
        L.orc_sample_mesh(len(pts), pts, len(fc), _ptr(fc), radius, seed, _ptr(P), _ptr(N), n)
    return P, N


def mesh_box(pts, nrm=None):
    """computeBoundingBox over removeNan's valid points (util.cpp:86-105,178-207)."""
    pts = f32(pts)
    ok = np.isfinite(pts).all(1)
    if nrm is not None:
        ok &= np.isfinite(f32(nrm)).all(1)
    v = pts[ok]
    return np.concatenate([v.min(0), v.max(0)]).astype(np.float32)


def train_mesh(pts, faces, nrm=None, rel=0.04, param=None, box6=None, vp=None, seed=0x5EED,
               ref_tail=1):
    """trainModel of a model with faces (ppf.cpp:46-70): the box of the given
    model, the mesh sampled at radius reSampleStep / 4, the sample trained on
    with that box."""
    param = param or default_train_param()
    box = f32(box6) if box6 is not None else None
    if box is None or np.float32(np.linalg.norm(box[3:] - box[:3])) == 0:
        box = mesh_box(pts, nrm)
    d = box[3:] - box[:3]
    D = np.sqrt(np.float32((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), dtype=np.float32)
    reStep = np.float32(D * np.float32(param.poseRefRelSamplingDistance))
    P, N = sample_mesh(pts, faces, np.float32(np.float64(reStep) / 4.0), seed)
    return train(P, N, rel=rel, param=param, box6=box, vp=vp, ref_tail=ref_tail)


def estimate_normals(pts, nrm, valid, query, k=10, smooth=1, invert=0, vp=None):
    pts = f32(pts).copy()
    nrm = f32(nrm).copy()
    vpa = None if vp is None else f32(vp)
    lib().orc_estimate_normals(len(pts), pts, nrm, np.ascontiguousarray(valid, np.int32),
                               len(valid), np.ascontiguousarray(query, np.int32), len(query), k,
                               smooth, invert, _ptr(vpa))
    return pts, nrm


def p2pl_system(src, T12, corr, dst, dstn, center, xi, tukey=0):
    out = np.zeros(28)
    lib().orc_p2pl_system(len(src), f32(src), np.ascontiguousarray(T12, np.float64),
                          np.ascontiguousarray(corr, np.int32), f32(dst), f32(dstn),
                          np.ascontiguousarray(center, np.float64),
                          np.ascontiguousarray(xi, np.float64), tukey, out)
    return out


def remove_nan(pts, nrm=None):
    """removeNan (util.cpp:86-105): ascending valid indices"""
    pts = f32(pts)
    nrm = None if nrm is None else f32(nrm)
    out = np.zeros(max(len(pts), 1), np.int32)
    m = lib().orc_remove_nan(len(pts), pts, _ptr(nrm), out)
    return out[:m]


def bounding_box(pts, idx):
    """computeBoundingBox (util.cpp:178-207) over idx: {min xyz, max xyz}"""
    pts = f32(pts)
    idx = np.ascontiguousarray(idx, np.int32)
    box = np.zeros(6, np.float32)
    lib().orc_bounding_box(len(pts), pts, idx if len(idx) else np.zeros(1, np.int32), len(idx),
                           box)
    return box


def train(pts, nrm=None, rel=0.04, param=None, box6=None, vp=None, ref_tail=1):
    param = param or default_train_param()
    m = Model()
    pts = f32(pts)
    nrm = None if nrm is None else f32(nrm)
    box = None if box6 is None else f32(box6)
    vpa = None if vp is None else f32(vp)
    rc = lib().orc_train(len(pts), pts, _ptr(nrm), _ptr(box), _ptr(vpa), rel, C.byref(param),
                         ref_tail, C.byref(m))
    if rc != 0:
        raise ValueError("orc_train failed")
    return m


def match(model: Model, pts, nrm=None, rel=0.04, kp_frac=0.1, min_score=0.2, param=None,
          box6=None, vp=None, cap=64, threads=1):
    param = param or default_match_param()
    pts = f32(pts)
    nrm = None if nrm is None else f32(nrm)
    box = None if box6 is None else f32(box6)
    vpa = None if vp is None else f32(vp)
    poses = np.zeros((cap, 12), np.float32)
    scores = np.zeros(cap, np.float32)
    st = MatchStats()
    n = lib().orc_match(C.byref(model), len(pts), pts, _ptr(nrm), _ptr(box), _ptr(vpa), rel,
                        kp_frac, min_score, C.byref(param), cap, poses, scores, C.byref(st),
                        threads)
    if n < 0:
        raise ValueError("orc_match failed")
    return poses[:n].reshape(-1, 3, 4), scores[:n], st


# ---- the CPU baseline's timing mode (oracle/cpu_ref.cpp, built into
# oracle/_ref by `make -C oracle ref` from the reference tree)
CPUREF_STAGES = ("scene remove nan", "scene kdtree", "scene sample1", "scene compute normal",
                 "scene sample2", "scene ppf", "cluster", "icp prepare", "icp", "total")


class CpuRefStats(C.Structure):
    _fields_ = [("ms", C.c_float * 10), ("nSampled", C.c_int), ("nKeypoints", C.c_int),
                ("nVotedKeypoints", C.c_int), ("nPoses", C.c_int), ("nResults", C.c_int),
                ("threads", C.c_int), ("nPairs", C.c_long), ("nVotes", C.c_long)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "ms"}
        d["stage_ms"] = {name: round(float(self.ms[i]), 2) for i, name in enumerate(CPUREF_STAGES)}
        return d


_cpuref = None


def cpuref_lib():
    global _cpuref
    if _cpuref is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libcpu_ref.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `make -C oracle ref`")
        L = C.CDLL(path)
        L.cpuref_model.restype = C.c_void_p
        L.cpuref_model.argtypes = [C.POINTER(Model)]
        L.cpuref_model_free.argtypes = [C.c_void_p]
        L.cpuref_match.restype = C.c_int
        L.cpuref_match.argtypes = [C.c_void_p, C.c_int, fp, C.c_void_p, C.c_float, C.c_float,
                                   C.c_float, C.POINTER(MatchParam), C.c_int, C.c_int, fp, fp,
                                   C.c_int, C.POINTER(CpuRefStats)]
        _cpuref = L
    return _cpuref


class CpuRefModel:
    """A trained oracle model prepared for the timing mode (hash map,
    frames); keeps the oracle model alive."""

    def __init__(self, model: Model):
        self.model = model
        self.h = cpuref_lib().cpuref_model(C.byref(model))

    def match(self, pts, rel=0.04, kp_frac=0.1, min_score=0.2, param=None, vp=None,
              threads=1, kp_stride=1, cap=64):
        param = param or default_match_param()
        pts = f32(pts)
        vpa = None if vp is None else f32(vp)
        poses = np.zeros((cap, 12), np.float32)
        scores = np.zeros(cap, np.float32)
        st = CpuRefStats()
        n = cpuref_lib().cpuref_match(self.h, len(pts), pts, _ptr(vpa), rel, kp_frac, min_score,
                                      C.byref(param), threads, kp_stride, poses, scores, cap,
                                      C.byref(st))
        k = min(n, cap)
        return poses[:k].reshape(-1, 3, 4), scores[:k], st.as_dict()

    def __del__(self):
        if getattr(self, "h", None):
            cpuref_lib().cpuref_model_free(self.h)
            self.h = None
