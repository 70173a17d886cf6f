"""Shared test setup: the `gpu` marker, import paths, seeded workloads."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def workload1():
    from surface_match_amd import synth
    return synth.workload(1)


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


@pytest.fixture(scope="session")
def H(tmp_path_factory):
    """canon.h compiled for the host (the device arithmetic, run on the CPU)."""
    so = str(tmp_path_factory.mktemp("canon") / "canon_host.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                           "-fno-fast-math", os.path.join(ROOT, "tests", "cpp", "canon_host.cpp"),
                           "-o", so])
    L = C.CDLL(so)
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.ch_atan2f.restype = C.c_float
    L.ch_atan2f.argtypes = [C.c_float, C.c_float]
    L.ch_ppf_key.restype = C.c_uint32
    L.ch_ppf_key.argtypes = [fp, fp, fp, fp, C.c_float, C.c_float, C.c_int]
    L.ch_frame.argtypes = [fp, fp, fp]
    L.ch_alpha.restype = C.c_float
    L.ch_alpha.argtypes = [fp, fp]
    L.ch_bin.restype = C.c_int
    L.ch_bin.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int]
    L.ch_bin_wrapped.restype = C.c_int
    L.ch_bin_wrapped.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int]
    L.ch_libm_divergence.argtypes = [fp, fp, C.c_int, C.c_float, C.c_float, C.c_int,
                                     np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]
    L.ch_bin_wrapped_fast.restype = C.c_int
    L.ch_bin_wrapped_fast.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int]
    L.ch_pose.argtypes = [fp, fp, C.c_float, C.c_float, fp]
    L.ch_quat.argtypes = [fp, fp]
    up = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    L.ch_philox.argtypes = [up, C.c_uint32, C.c_uint32]
    L.ch_uniform01.restype = C.c_float
    L.ch_uniform01.argtypes = [C.c_uint32]
    L.ch_mesh_point.argtypes = [fp, fp, fp, C.c_float, C.c_float, fp]
    L.ch_expf.restype = C.c_float
    L.ch_expf.argtypes = [C.c_float]
    L.ch_unit_orthogonal.argtypes = [fp, fp]
    return L
