"""The C ABI library: loads, exports every declared entry point, and its
host-side parameter logic matches the reference's constructors.  CPU only
(no call here launches a kernel)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def N():
    from surface_match_amd import _native
    return _native


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ppf_c.h")).read()
    return sorted(set(re.findall(r"PPF_API\s+[^;(]*?\b(ppf_\w+)\s*\(", src)))


def test_exports_every_declared_symbol(N):
    L = N.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(N.EXPORTS) == names


def test_train_param_defaults_and_checks(N):
    L = N.lib()
    p = N.TrainParam()
    assert L.ppf_train_param_default(C.byref(p)) == 0
    assert (round(p.featDistanceStepRel, 6), p.featAngleResolution,
            round(p.poseRefRelSamplingDistance, 6), p.knnNormal, p.smoothNormal) == (0.04, 30, 0.01, 10, 1)
    assert L.ppf_train_param_check(C.byref(p)) == 0
    for field, bad, msg in [("featDistanceStepRel", 0.0, "featDistanceStepRel"),
                            ("featDistanceStepRel", 1.5, "featDistanceStepRel"),
                            ("featAngleResolution", 0, "featAngleResolution"),
                            ("poseRefRelSamplingDistance", 0.0, "poseRefRelSamplingDistance")]:
        q = N.TrainParam()
        L.ppf_train_param_default(C.byref(q))
        setattr(q, field, bad)
        assert L.ppf_train_param_check(C.byref(q)) == N.PPF_ERR_RANGE
        assert msg in L.ppf_last_error().decode()


def test_match_param_checks_keep_reference_guard(N):
    """type.cpp:115: the step/threshold checks only run when dense
    refinement is OFF (the reference's inverted guard)."""
    L = N.lib()
    p = N.MatchParam()
    assert L.ppf_match_param_default(C.byref(p)) == 0
    assert p.numMatches == 1 and p.densePoseRefinement == 1 and p.poseRefNumSteps == 5
    p.poseRefNumSteps = 0
    assert L.ppf_match_param_check(C.byref(p)) == 0          # guard skips the check
    p.densePoseRefinement = 0
    assert L.ppf_match_param_check(C.byref(p)) == N.PPF_ERR_RANGE
    assert "poseRefNumSteps" in L.ppf_last_error().decode()
    p = N.MatchParam()
    L.ppf_match_param_default(C.byref(p))
    p.numMatches = 0
    assert L.ppf_match_param_check(C.byref(p)) == N.PPF_ERR_RANGE
    p.numMatches, p.maxOverlapDistRel, p.maxOverlapDistAbs = 1, -1.0, -1.0
    assert L.ppf_match_param_check(C.byref(p)) == N.PPF_ERR_RANGE


def test_python_mirror_raises_range_error(N):
    from surface_match_amd import TrainParam, MatchParam
    with pytest.raises(ValueError, match="featAngleResolution"):
        TrainParam(featAngleResolution=0)
    with pytest.raises(ValueError, match="numMatches"):
        MatchParam(numMatches=0)


def test_null_arguments_rejected(N):
    L = N.lib()
    assert L.ppf_detector_create(None) == N.PPF_ERR_ARG
    assert L.ppf_train_model(None, None, None, 0, None, None, 0.04, None) == N.PPF_ERR_ARG
    assert L.ppf_save(None, b"x") == N.PPF_ERR_ARG


def test_no_device_fails_loudly(N):
    """Without a GPU the library reports a device error; it has no CPU path."""
    L = N.lib()
    h = C.c_void_p()
    rc = L.ppf_detector_create(C.byref(h))
    if rc == 0:
        L.ppf_detector_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == N.PPF_ERR_DEVICE
    assert "HIP" in L.ppf_last_error().decode()
