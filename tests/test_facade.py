"""The C++ drop-in facade (include/ppf/ppf.h): compiles against the C ABI
without Eigen (CPU) and drives train -> save -> load -> match on the GPU."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
LIBDIR = os.path.join(ROOT, "surface_match_amd", "_lib")


def _compile(out):
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), SRC,
                           "-L" + LIBDIR, "-lppf_amd", "-Wl,-rpath," + LIBDIR, "-o", out])


def test_facade_compiles_and_links(tmp_path):
    _compile(str(tmp_path / "facade"))


def _write(path, pts, nrm=None):
    with open(path, "wb") as f:
        np.array([len(pts)], np.int32).tofile(f)
        np.ascontiguousarray(pts, np.float32).tofile(f)
        if nrm is not None:
            np.ascontiguousarray(nrm, np.float32).tofile(f)


@pytest.mark.gpu
def test_facade_end_to_end(tmp_path):
    from surface_match_amd import synth
    w = synth.workload(1, scale=0.5)
    mp, mn = w["model"]
    _write(tmp_path / "m.bin", mp, mn)
    _write(tmp_path / "s.bin", w["scene"].points)
    exe = str(tmp_path / "facade")
    _compile(exe)
    out = subprocess.run([exe, str(tmp_path / "m.bin"), str(tmp_path / "s.bin"),
                          str(tmp_path / "x.model")], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "No trained model in matchScene" in out.stdout
    assert "featDistanceStepRel" in out.stdout
    lines = [l for l in out.stdout.splitlines() if l.startswith("score")]
    assert len(lines) >= 2
    D = 147.3
    for l in lines:
        vals = [float(v) for v in l.split()[3:]]
        T = np.array(vals, np.float64).reshape(4, 4).T   # column-major
        k, ang, dt = synth.pose_error(T, w["scene"].gt_poses, D)
        assert ang < 1.0 and dt < 0.01 * D
