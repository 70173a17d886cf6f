"""ctypes binding of libppf_amd.so (the C ABI in include/ppf_c.h).

The library must have been built (``python -m surface_match_amd.build``);
importing this module never falls back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libppf_amd.so")

PPF_OK, PPF_ERR_RANGE, PPF_ERR_RUNTIME, PPF_ERR_DEVICE, PPF_ERR_ARG = range(5)


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


class MatchResultC(C.Structure):
    _fields_ = [("sampledPoints", C.c_void_p), ("sampledNormals", C.c_void_p),
                ("sampledCap", C.c_size_t), ("nSampled", C.c_size_t),
                ("keyPoints", C.c_void_p), ("keyNormals", C.c_void_p),
                ("keyCap", C.c_size_t), ("nKey", C.c_size_t)]


class MatchStats(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nScene", "nValid", "nSampled", "nKeypoints",
                                       "nVotedKeypoints", "nPoses", "nClusters", "nHypotheses",
                                       "nSparseIcp", "nDenseIcp", "nResults", "nDense")] + \
               [(n, C.c_ulonglong) for n in ("nPairs", "nKeyHits", "nVotes")] + \
               [(n, C.c_float) for n in ("msSample", "msNormals", "msKeypoints", "msVote",
                                         "msCluster", "msIcpPrepare", "msIcp", "msTotal",
                                         "msVoteKernel", "msIcpSparse", "msIcpDense",
                                         "msPairKernel", "msHitsKernel")] + \
               [("nVoteFallback", C.c_int), ("msIngest", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ModelInfo(C.Structure):
    _fields_ = [("nSampled", C.c_int), ("nReSampled", C.c_int), ("nKeys", C.c_int),
                ("nEntries", C.c_int), ("angleNum", C.c_int), ("diameter", C.c_float),
                ("samplingDistanceRel", C.c_float), ("boxMin", C.c_float * 3),
                ("boxMax", C.c_float * 3)]


class IcpCriteria(C.Structure):
    _fields_ = [("iterations", C.c_int), ("rejectDist", C.c_float), ("mseMin", C.c_float),
                ("mseMax", C.c_float), ("tolerance", C.c_float)]


class IcpResult(C.Structure):
    _fields_ = [("pose16", C.c_float * 16), ("type", C.c_int), ("mse", C.c_float),
                ("convergeRate", C.c_float), ("iterations", C.c_int), ("converged", C.c_int),
                ("inliers", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("voteBlock", "voteXcdRemap", "voteCellDiv", "icpBlock",
                                       "icpSpread", "normCell4", "voteSplit",
                                       "voteHitCap", "voteTiles", "voteOrder")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MlsStats(C.Structure):
    _fields_ = [("neighbours", C.c_longlong), ("queries", C.c_int), ("deviceMs", C.c_float)]


class TuneReport(C.Structure):
    _fields_ = [("best", Tuning), ("voteMs", C.c_float), ("icpMs", C.c_float),
                ("configsTried", C.c_int), ("normalsMs", C.c_float)]


_lib = None

# every function declared in include/ppf_c.h
EXPORTS = [
    "ppf_train_param_default", "ppf_match_param_default", "ppf_train_param_check",
    "ppf_match_param_check", "ppf_detector_create", "ppf_detector_destroy", "ppf_train_model",
    "ppf_match_scene", "ppf_save", "ppf_load", "ppf_last_error", "ppf_scene_upload",
    "ppf_scene_free", "ppf_match_scene_resident", "ppf_model_get_info", "ppf_model_get_sampled",
    "ppf_model_get_table", "ppf_pair_features", "ppf_vote", "ppf_voxel_sample",
    "ppf_radius_search", "ppf_knn_search", "ppf_estimate_normals", "ppf_device_math",
    "ppf_icp_target_create", "ppf_icp_target_destroy", "ppf_icp_refine", "ppf_icp_adjoint",
    "ppf_autotune", "ppf_set_tuning", "ppf_get_tuning", "ppf_match_scenes_resident",
    "ppf_scene_inliers", "ppf_measure_ceilings", "ppf_ply_read", "ppf_ply_info", "ppf_ply_copy",
    "ppf_ply_free", "ppf_ply_write", "ppf_scene_upload_ply", "ppf_train_model_mesh",
    "ppf_sample_mesh", "ppf_estimate_normals_mls", "ppf_scene_download",
    "ppf_voxel_sample_noalloc", "ppf_icp_refine_multi",
]


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m surface_match_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, sz, f, i = C.c_void_p, C.c_size_t, C.c_float, C.c_int
    sig = {
        "ppf_train_param_default": (i, [C.POINTER(TrainParam)]),
        "ppf_match_param_default": (i, [C.POINTER(MatchParam)]),
        "ppf_train_param_check": (i, [C.POINTER(TrainParam)]),
        "ppf_match_param_check": (i, [C.POINTER(MatchParam)]),
        "ppf_detector_create": (i, [C.POINTER(vp)]),
        "ppf_detector_destroy": (None, [vp]),
        "ppf_train_model": (i, [vp, vp, vp, sz, vp, vp, f, C.POINTER(TrainParam)]),
        "ppf_match_scene": (i, [vp, vp, vp, sz, vp, vp, f, f, f, C.POINTER(MatchParam), vp, vp,
                                sz, C.POINTER(sz), vp, vp]),
        "ppf_save": (i, [vp, C.c_char_p]),
        "ppf_load": (i, [vp, C.c_char_p]),
        "ppf_last_error": (C.c_char_p, []),
        "ppf_scene_upload": (i, [vp, vp, vp, sz, vp, vp, C.POINTER(vp)]),
        "ppf_scene_free": (None, [vp]),
        "ppf_scene_download": (i, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(i), vp, vp, vp,
                                   vp]),
        "ppf_match_scene_resident": (i, [vp, vp, f, f, f, C.POINTER(MatchParam), vp, vp, sz,
                                         C.POINTER(sz), vp]),
        "ppf_model_get_info": (i, [vp, C.POINTER(ModelInfo)]),
        "ppf_model_get_sampled": (i, [vp, vp, vp, vp, vp]),
        "ppf_model_get_table": (i, [vp, vp, vp, vp, vp]),
        "ppf_pair_features": (i, [vp, vp, vp, sz, f, f, i, vp, vp]),
        "ppf_vote": (i, [vp, vp, vp, sz, vp, sz, f, vp, vp, sz, vp, vp]),
        "ppf_voxel_sample": (i, [vp, sz, vp, sz, f, vp, vp, C.POINTER(sz)]),
        "ppf_voxel_sample_noalloc": (i, [vp, vp, vp, sz, vp, sz, f, vp, vp, vp, vp]),
        "ppf_radius_search": (i, [vp, sz, vp, sz, f, f, vp, vp, sz]),
        "ppf_knn_search": (i, [vp, sz, vp, sz, i, f, vp, vp]),
        "ppf_estimate_normals": (i, [vp, vp, sz, vp, sz, vp, sz, i, i, i, vp]),
        "ppf_device_math": (i, [i, vp, vp, sz, vp]),
        "ppf_icp_target_create": (i, [vp, vp, sz, f, C.POINTER(vp)]),
        "ppf_icp_target_destroy": (None, [vp]),
        "ppf_icp_refine": (i, [vp, vp, vp, sz, vp, sz, C.POINTER(IcpCriteria), i, i, i, f,
                               vp, C.POINTER(f)]),
        "ppf_icp_adjoint": (i, [vp, vp, sz, vp, vp, sz, f, i, vp, vp]),
        "ppf_icp_refine_multi": (i, [vp, sz, vp, vp, vp, sz, vp, sz, C.POINTER(IcpCriteria), i, i,
                                     i, f, vp, C.POINTER(f)]),
        "ppf_autotune": (i, [vp, vp, f, f, f, C.POINTER(MatchParam), C.POINTER(TuneReport)]),
        "ppf_set_tuning": (i, [vp, C.POINTER(Tuning)]),
        "ppf_get_tuning": (i, [vp, C.POINTER(Tuning)]),
        "ppf_match_scenes_resident": (i, [vp, vp, sz, f, f, f, C.POINTER(MatchParam), vp, vp, sz,
                                          vp, C.POINTER(MatchStats)]),
        "ppf_scene_inliers": (i, [vp, vp, vp, sz, i, f, vp]),
        "ppf_measure_ceilings": (i, [C.POINTER(f), C.POINTER(f)]),
        "ppf_ply_read": (i, [C.c_char_p, C.POINTER(vp)]),
        "ppf_ply_info": (i, [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]),
        "ppf_ply_copy": (i, [vp, vp, vp, vp]),
        "ppf_ply_free": (None, [vp]),
        "ppf_ply_write": (i, [C.c_char_p, vp, vp, sz, vp, sz, i]),
        "ppf_scene_upload_ply": (i, [vp, C.c_char_p, vp, C.POINTER(vp)]),
        "ppf_train_model_mesh": (i, [vp, vp, vp, sz, vp, sz, vp, vp, f, C.POINTER(TrainParam),
                                     C.c_uint64]),
        "ppf_sample_mesh": (i, [vp, sz, vp, sz, f, C.c_uint64, vp, vp, sz, C.POINTER(sz)]),
        "ppf_estimate_normals_mls": (i, [vp, vp, sz, vp, sz, vp, sz, f, i, i, vp,
                                         C.POINTER(MlsStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class PPFError(RuntimeError):
    pass


def check(rc: int):
    """Raise the exception class the reference raises for this status."""
    if rc == PPF_OK:
        return
    msg = lib().ppf_last_error().decode()
    if rc == PPF_ERR_RANGE:
        raise ValueError(msg)          # std::range_error
    if rc == PPF_ERR_RUNTIME:
        raise RuntimeError(msg)        # std::runtime_error
    raise PPFError(msg)


def ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)
