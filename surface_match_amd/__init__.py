"""MI355X-native PPF detection path, drop-in for surface_match's ppf::Detector.

Python mirror of the reference's public API (ppf.h:10-62, type.h:8-181):
``Detector.trainModel / matchScene / save / load`` with ``PointCloud``,
``TrainParam``, ``MatchParam`` and ``MatchResult``.  Every call goes
through the C ABI of libppf_amd.so (include/ppf_c.h) into gfx950 HIP
kernels; nothing here computes on the CPU beyond argument marshalling.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

__all__ = ["PointCloud", "TrainParam", "MatchParam", "MatchResult", "Detector", "BoundingBox",
           "ResidentScene", "readPLY", "writePLY", "sampleMesh", "estimateNormalMLS"]


@dataclass
class BoundingBox:
    """type.h:8-20"""
    min: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    max: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))

    def size(self):
        return np.asarray(self.max, np.float32) - np.asarray(self.min, np.float32)

    def center(self):
        return (np.asarray(self.min, np.float32) + np.asarray(self.max, np.float32)) / np.float32(2)

    def diameter(self) -> float:
        s = self.size()
        return float(np.sqrt(np.float32(s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]))


@dataclass
class PointCloud:
    """type.h:22-33: points / normals are n x 3 float arrays; normals are
    present when their count equals the point count (type.cpp:36-38)."""
    point: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    normal: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.float32))
    face: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    box: BoundingBox = field(default_factory=BoundingBox)
    viewPoint: np.ndarray = field(default_factory=lambda: np.full(3, np.nan, np.float32))

    def hasNormal(self) -> bool:
        return len(self.point) > 0 and len(self.point) == len(self.normal)

    def size(self) -> int:
        return len(self.point)

    def empty(self) -> bool:
        return len(self.point) == 0

    def _c_args(self):
        p = N.f32(np.reshape(self.point, (-1, 3)))
        n = N.f32(np.reshape(self.normal, (-1, 3))) if self.hasNormal() else None
        box = N.f32(np.concatenate([np.asarray(self.box.min, np.float32).ravel(),
                                    np.asarray(self.box.max, np.float32).ravel()]))
        vp = N.f32(np.asarray(self.viewPoint, np.float32).ravel())
        return p, n, box, vp


def TrainParam(featDistanceStepRel=0.04, featAngleResolution=30, poseRefRelSamplingDistance=0.01,
               knnNormal=10, smoothNormal=True) -> N.TrainParam:
    """type.h:71-73 / type.cpp:71-89 (raises ValueError = std::range_error)."""
    p = N.TrainParam(featDistanceStepRel, featAngleResolution, poseRefRelSamplingDistance,
                     knnNormal, int(bool(smoothNormal)))
    N.check(N.lib().ppf_train_param_check(C.byref(p)))
    return p


def MatchParam(numMatches=1, knnNormal=10, smoothNormal=True, invertNormal=False,
               maxOverlapDistRel=0.5, maxOverlapDistAbs=0.0, sparsePoseRefinement=True,
               densePoseRefinement=True, poseRefNumSteps=5, poseRefDistThresholdRel=0.1,
               poseRefDistThresholdAbs=0.0, poseRefScoringDistRel=0.01,
               poseRefScoringDistAbs=0.0) -> N.MatchParam:
    """type.h:170-175 / type.cpp:91-127."""
    p = N.MatchParam(numMatches, knnNormal, int(bool(smoothNormal)), int(bool(invertNormal)),
                     maxOverlapDistRel, maxOverlapDistAbs, int(bool(sparsePoseRefinement)),
                     int(bool(densePoseRefinement)), poseRefNumSteps, poseRefDistThresholdRel,
                     poseRefDistThresholdAbs, poseRefScoringDistRel, poseRefScoringDistAbs)
    N.check(N.lib().ppf_match_param_check(C.byref(p)))
    return p


def readPLY(filename: str, mesh: PointCloud) -> bool:
    """filePLY.h:6: fills mesh.point / normal / face from a PLY file; False
    (after printing the reason, as the reference does) when it cannot."""
    h = C.c_void_p()
    rc = N.lib().ppf_ply_read(str(filename).encode(), C.byref(h))
    if rc != 0:
        print(N.lib().ppf_last_error().decode())
        return False
    try:
        nv, nn, nf = C.c_size_t(), C.c_size_t(), C.c_size_t()
        N.check(N.lib().ppf_ply_info(h, C.byref(nv), C.byref(nn), C.byref(nf)))
        pts = np.zeros((nv.value, 3), np.float32)
        nrm = np.zeros((nn.value, 3), np.float32)
        fac = np.zeros((nf.value, 3), np.int32)
        N.check(N.lib().ppf_ply_copy(h, N.ptr(pts), N.ptr(nrm), N.ptr(fac)))
    finally:
        N.lib().ppf_ply_free(h)
    mesh.point, mesh.normal, mesh.face = pts, nrm, fac
    return True


def writePLY(filename: str, mesh: PointCloud, write_ascii: bool = False) -> bool:
    """filePLY.h:8-9: binary little endian (or ASCII) PLY of the points,
    normals when present, and triangles."""
    p = N.f32(np.reshape(mesh.point, (-1, 3)))
    n = N.f32(np.reshape(mesh.normal, (-1, 3))) if mesh.hasNormal() else None
    f = N.i32(np.reshape(mesh.face, (-1, 3)))
    rc = N.lib().ppf_ply_write(str(filename).encode(), N.ptr(p), N.ptr(n), len(p), N.ptr(f),
                               len(f), int(bool(write_ascii)))
    if rc != 0:
        print(N.lib().ppf_last_error().decode())
        return False
    return True


def sampleMesh(pc: PointCloud, radius: float, seed: int = 0x5EED) -> PointCloud:
    """util.h:8 / util.cpp:20-84: int(area / radius^2) points spread over the
    triangles by area, each with its triangle's unit normal; the box is
    carried over.  The uniforms come from Philox4x32-10 keyed by seed (the
    reference seeds from std::random_device).  No faces: an empty cloud."""
    p = N.f32(np.reshape(pc.point, (-1, 3)))
    f = N.i32(np.reshape(pc.face, (-1, 3)))
    out = PointCloud()
    if len(f) == 0:
        return out
    n = C.c_size_t()
    L = N.lib()
    N.check(L.ppf_sample_mesh(N.ptr(p), len(p), N.ptr(f), len(f), radius, seed, None, None, 0,
                              C.byref(n)))
    P = np.zeros((n.value, 3), np.float32)
    Q = np.zeros((n.value, 3), np.float32)
    if n.value:
        N.check(L.ppf_sample_mesh(N.ptr(p), len(p), N.ptr(f), len(f), radius, seed, N.ptr(P),
                                  N.ptr(Q), n.value, C.byref(n)))
    out.point, out.normal, out.box = P, Q, pc.box
    return out


def estimateNormalMLS(pc: PointCloud, indices, radius: float, order: int = 2,
                      invert: bool = False, validIndices=None) -> dict:
    """estimateNormalMLS (privateUtil.h:13-14, util.cpp:336-457) on the GPU:
    the points listed in `indices` move onto a weighted polynomial fit
    (order 1..3) of their neighbours within `radius` and take its normal,
    oriented towards pc.viewPoint (or +z).  The reference's kd-tree holds
    the cloud's valid points; here `validIndices` (default: removeNan).
    Updates pc.point / pc.normal in place (normals are created as NaN when
    absent); returns the neighbour total and the device time."""
    p = N.f32(np.reshape(pc.point, (-1, 3))).copy()
    n = (N.f32(np.reshape(pc.normal, (-1, 3))).copy() if pc.hasNormal()
         else np.full_like(p, np.nan))
    if validIndices is None:
        validIndices = np.flatnonzero(np.isfinite(p).all(1))
    v = N.i32(validIndices)
    q = N.i32(indices)
    vp = N.f32(np.asarray(pc.viewPoint, np.float32).ravel())
    st = N.MlsStats()
    N.check(N.lib().ppf_estimate_normals_mls(N.ptr(p), N.ptr(n), len(p), N.ptr(v), len(v),
                                             N.ptr(q), len(q), radius, order, int(bool(invert)),
                                             N.ptr(vp), C.byref(st)))
    pc.point, pc.normal = p, n
    return {"neighbours": int(st.neighbours), "queries": int(st.queries),
            "device_ms": float(st.deviceMs)}


@dataclass
class MatchResult:
    """type.h:178-181"""
    sampledScene: PointCloud = field(default_factory=PointCloud)
    keyPoint: PointCloud = field(default_factory=PointCloud)


class Detector:
    """ppf::Detector (ppf.h:10-62) backed by libppf_amd.so."""

    def __init__(self):
        h = C.c_void_p()
        N.check(N.lib().ppf_detector_create(C.byref(h)))
        self._h = h
        self.last_stats: dict | None = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().ppf_detector_destroy(h)
            self._h = None

    def trainModel(self, model: PointCloud, samplingDistanceRel: float = 0.04,
                   param: N.TrainParam | None = None, mesh_seed: int = 0x5EED):
        """ppf.cpp:36-147.  A model with faces is sampled first (sampleMesh,
        seeded by mesh_seed where the reference draws from random_device)."""
        p, n, box, vp = model._c_args()
        param = param or TrainParam()
        f = N.i32(np.reshape(model.face, (-1, 3)))
        if len(f):
            N.check(N.lib().ppf_train_model_mesh(self._h, N.ptr(p), N.ptr(n), len(p), N.ptr(f),
                                                 len(f), N.ptr(box), N.ptr(vp),
                                                 samplingDistanceRel, C.byref(param), mesh_seed))
        else:
            N.check(N.lib().ppf_train_model(self._h, N.ptr(p), N.ptr(n), len(p), N.ptr(box),
                                            N.ptr(vp), samplingDistanceRel, C.byref(param)))

    def matchScene(self, scene: PointCloud, samplingDistanceRel: float = 0.04,
                   keyPointFraction: float = 0.2, minScore: float = 0.2,
                   param: N.MatchParam | None = None, matchResult: MatchResult | None = None):
        """Returns (poses [k,4,4] float32, scores [k]) like the reference's
        output vectors (ppf.cpp:501-507)."""
        p, n, box, vp = scene._c_args()
        param = param or MatchParam()
        cap = max(int(param.numMatches), 1)
        poses = np.zeros((cap, 16), np.float32)
        scores = np.zeros(cap, np.float32)
        nout = C.c_size_t(0)
        st = N.MatchStats()
        mr = None
        bufs = []
        if matchResult is not None:
            cap_s = len(p)
            sp, sn = np.zeros((cap_s, 3), np.float32), np.zeros((cap_s, 3), np.float32)
            kp, kn = np.zeros((cap_s, 3), np.float32), np.zeros((cap_s, 3), np.float32)
            bufs = [sp, sn, kp, kn]
            mr = N.MatchResultC(N.ptr(sp), N.ptr(sn), cap_s, 0, N.ptr(kp), N.ptr(kn), cap_s, 0)
        N.check(N.lib().ppf_match_scene(self._h, N.ptr(p), N.ptr(n), len(p), N.ptr(box), N.ptr(vp),
                                        samplingDistanceRel, keyPointFraction, minScore,
                                        C.byref(param), N.ptr(poses), N.ptr(scores), cap,
                                        C.byref(nout), C.byref(mr) if mr else None, C.byref(st)))
        self.last_stats = st.as_dict()
        if matchResult is not None:
            sp, sn, kp, kn = bufs
            matchResult.sampledScene = PointCloud(sp[:mr.nSampled].copy(), sn[:mr.nSampled].copy())
            matchResult.keyPoint = PointCloud(kp[:mr.nKey].copy(), kn[:mr.nKey].copy())
        k = nout.value
        # column-major 4x4 -> row-major numpy
        return poses[:k].reshape(k, 4, 4).transpose(0, 2, 1).copy(), scores[:k].copy()

    def get_tuning(self) -> dict:
        """Launch shapes in use (vote block / XCD mapping / grid cell, ICP block / spread)."""
        t = N.Tuning()
        N.check(N.lib().ppf_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Change launch shapes; results do not depend on them."""
        t = N.Tuning()
        N.check(N.lib().ppf_get_tuning(self._h, C.byref(t)))
        for k, v in kw.items():
            if k not in t.as_dict():
                raise KeyError(k)
            setattr(t, k, int(v))
        N.check(N.lib().ppf_set_tuning(self._h, C.byref(t)))

    def save(self, filename: str):
        N.check(N.lib().ppf_save(self._h, str(filename).encode()))

    def load(self, filename: str):
        N.check(N.lib().ppf_load(self._h, str(filename).encode()))

    # -- introspection of the trained model (privateType.h:56-66)
    def model_info(self) -> N.ModelInfo:
        info = N.ModelInfo()
        N.check(N.lib().ppf_model_get_info(self._h, C.byref(info)))
        return info

    def sampled_model(self):
        info = self.model_info()
        sp = np.zeros((info.nSampled, 3), np.float32)
        sn = np.zeros((info.nSampled, 3), np.float32)
        rp = np.zeros((info.nReSampled, 3), np.float32)
        rn = np.zeros((info.nReSampled, 3), np.float32)
        N.check(N.lib().ppf_model_get_sampled(self._h, N.ptr(sp), N.ptr(sn), N.ptr(rp), N.ptr(rn)))
        return sp, sn, rp, rn

    def table(self):
        info = self.model_info()
        keys = np.zeros(max(info.nKeys, 1), np.uint32)
        offs = np.zeros(info.nKeys + 1, np.int32)
        refs = np.zeros(max(info.nEntries, 1), np.uint32)
        al = np.zeros(max(info.nEntries, 1), np.float32)
        N.check(N.lib().ppf_model_get_table(self._h, N.ptr(keys), N.ptr(offs), N.ptr(refs),
                                            N.ptr(al)))
        return keys[:info.nKeys], offs, refs[:info.nEntries], al[:info.nEntries]

    def vote(self, pts, nrm, keypoints, minScore=0.2, dump=0):
        """The vote stage alone over a given sampled scene (ppf.cpp:247-409):
        returns (candidates [nk,3,3] (ref, bin, vote; -1 = none), poses [np,3,4],
        accumulator bins of the first `dump` key points)."""
        pts, nrm, kp = N.f32(pts), N.f32(nrm), N.i32(keypoints)
        info = self.model_info()
        cand = np.zeros((len(kp), 3, 3), np.int32)
        acc = np.zeros((max(dump, 1), info.nSampled, info.angleNum), np.int32)
        poses = np.zeros((len(kp) * 3, 12), np.float32)
        npo = C.c_int(0)
        N.check(N.lib().ppf_vote(self._h, N.ptr(pts), N.ptr(nrm), len(pts), N.ptr(kp), len(kp),
                                 minScore, N.ptr(cand), N.ptr(acc) if dump else None, dump,
                                 N.ptr(poses), C.byref(npo)))
        return cand, poses[:npo.value].reshape(-1, 3, 4), acc[:dump]


class ResidentScene:
    """A scene kept in HBM for repeated matches (ppf_scene_upload)."""

    def __init__(self, det: Detector, scene: PointCloud | None = None, ply: str | None = None,
                 viewPoint=None):
        h = C.c_void_p()
        if ply is not None:
            vp = None if viewPoint is None else N.f32(np.asarray(viewPoint, np.float32).ravel())
            N.check(N.lib().ppf_scene_upload_ply(det._h, str(ply).encode(), N.ptr(vp),
                                                 C.byref(h)))
        else:
            p, n, box, vp = scene._c_args()
            N.check(N.lib().ppf_scene_upload(det._h, N.ptr(p), N.ptr(n), len(p), N.ptr(box),
                                             N.ptr(vp), C.byref(h)))
        self._h = h
        self._det = det

    def match(self, samplingDistanceRel=0.04, keyPointFraction=0.2, minScore=0.2, param=None):
        param = param or MatchParam()
        cap = max(int(param.numMatches), 1)
        poses = np.zeros((cap, 16), np.float32)
        scores = np.zeros(cap, np.float32)
        nout = C.c_size_t(0)
        st = N.MatchStats()
        N.check(N.lib().ppf_match_scene_resident(self._det._h, self._h, samplingDistanceRel,
                                                 keyPointFraction, minScore, C.byref(param),
                                                 N.ptr(poses), N.ptr(scores), cap, C.byref(nout),
                                                 C.byref(st)))
        k = nout.value
        return poses[:k].reshape(k, 4, 4).transpose(0, 2, 1).copy(), scores[:k].copy(), st.as_dict()

    def download(self):
        """What the upload kept: (points [n,3], normals [n,3] or None, valid
        indices, box [6]) -- the cloud as given, removeNan's indices and the
        box the match uses (ppf.cpp:153,170-176)."""
        n, nv, hn = C.c_size_t(), C.c_size_t(), C.c_int()
        N.check(N.lib().ppf_scene_download(self._h, C.byref(n), C.byref(nv), C.byref(hn), None,
                                           None, None, None))
        pts = np.zeros((n.value, 3), np.float32)
        nrm = np.zeros((n.value, 3), np.float32) if hn.value else None
        valid = np.zeros(nv.value, np.int32)
        box = np.zeros(6, np.float32)
        N.check(N.lib().ppf_scene_download(self._h, None, None, None, N.ptr(pts), N.ptr(nrm),
                                           N.ptr(valid), N.ptr(box)))
        return pts, nrm, valid, box

    def inliers(self, poses, dense: bool, dist: float) -> np.ndarray:
        """inliner (util.cpp:628-641) of the sampled (dense=False) or
        resampled (dense=True) model at each pose (k x 4 x 4, row-major)
        against this scene as the last match left it."""
        P = np.ascontiguousarray(np.asarray(poses, np.float32).reshape(-1, 4, 4).transpose(0, 2, 1))
        out = np.zeros(len(P), np.int32)
        N.check(N.lib().ppf_scene_inliers(self._det._h, self._h, N.ptr(P), len(P), int(dense),
                                          dist, N.ptr(out)))
        return out

    @staticmethod
    def match_many(scenes, samplingDistanceRel=0.04, keyPointFraction=0.2, minScore=0.2,
                   param=None):
        """Match several resident scenes of one detector together
        (ppf_match_scenes_resident): returns [(poses [k,4,4], scores [k])] per
        scene, and the summed stats."""
        if not scenes:
            return [], {}
        det = scenes[0]._det
        if any(sc._det is not det for sc in scenes):
            raise ValueError("all scenes must belong to one detector")
        param = param or MatchParam()
        cap = max(int(param.numMatches), 1)
        n = len(scenes)
        poses = np.zeros((n, cap, 16), np.float32)
        scores = np.zeros((n, cap), np.float32)
        nout = np.zeros(n, np.uint64)
        hs = (C.c_void_p * n)(*[sc._h.value for sc in scenes])
        st = N.MatchStats()
        N.check(N.lib().ppf_match_scenes_resident(det._h, C.cast(hs, C.c_void_p), n,
                                                  samplingDistanceRel, keyPointFraction, minScore,
                                                  C.byref(param), N.ptr(poses), N.ptr(scores), cap,
                                                  N.ptr(nout), C.byref(st)))
        out = []
        for i in range(n):
            k = int(nout[i])
            out.append((poses[i, :k].reshape(k, 4, 4).transpose(0, 2, 1).copy(), scores[i, :k].copy()))
        return out, st.as_dict()

    def autotune(self, samplingDistanceRel=0.04, keyPointFraction=0.2, minScore=0.2, param=None):
        """Pick the vote/ICP launch shapes that run fastest on this scene."""
        param = param or MatchParam()
        rep = N.TuneReport()
        N.check(N.lib().ppf_autotune(self._det._h, self._h, samplingDistanceRel, keyPointFraction,
                                     minScore, C.byref(param), C.byref(rep)))
        return {"best": rep.best.as_dict(), "voteMs": rep.voteMs, "icpMs": rep.icpMs,
                "normalsMs": rep.normalsMs, "configsTried": rep.configsTried}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib().ppf_scene_free(h)
            self._h = None
