"""The C++ drop-in facade (include/ppf/ppf.h): compiles against the C ABI
without Eigen (CPU) and drives the reference example's sequence on the GPU:
readPLY -> train -> save -> load -> match -> transformPointCloud / writePLY,
then sampleMesh of a mesh file."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
LIBDIR = os.path.join(ROOT, "surface_match_amd", "_lib")


def _compile(out, src=SRC):
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src,
                           "-L" + LIBDIR, "-lppf_amd", "-Wl,-rpath," + LIBDIR, "-o", out])


def test_facade_compiles_and_links(tmp_path):
    _compile(str(tmp_path / "facade"))


def test_facade_host_helpers(tmp_path):
    """removeNan, computeBoundingBox, extraIndices, normalizeNormal,
    transformPointCloud and Timer (util.cpp:86-232, helper.h) on hand-made
    clouds, against values worked out by hand; no GPU call."""
    exe = str(tmp_path / "util")
    _compile(exe, os.path.join(ROOT, "tests", "cpp", "facade_util.cpp"))
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "removeNan 3 0 2 3 | 2 0 3"
    # box over points 0, 2, 3: x in [-4, 2], y in [-1, 5], z in [0.5, 7]
    assert lines[1].startswith("box -4 -1 0.5 2 5 7 diameter ")
    assert abs(float(lines[1].split()[-1]) - np.sqrt(36 + 36 + 6.5 ** 2)) < 1e-5
    # (2,-1,7) with the zero normal kept as is; (0,0,2) normalized; box of the two
    assert lines[2] == "extra 2 2 -1 7 n 0 0 0 | 0 0 1 box 1 7"
    assert lines[3] == "invert -0 -0 -1 | -0 -0 -0"
    # Rz(90): (x, y, z) -> (-y, x, z), then + (10, 20, 30)
    assert lines[4] == "transform 8 21 33 | 10 19 34 n 0 1 0 faces 1 box 8 21"
    assert lines[5] == "no normals 0"
    assert lines[6].startswith("timer cost(ms): ")
    assert lines[7] == "sampleMesh without faces 0"

@pytest.mark.gpu
def test_facade_end_to_end(tmp_path):
    from surface_match_amd import PointCloud, readPLY, writePLY, synth
    w = synth.workload(1, scale=0.5)
    mp, mn = w["model"]
    assert writePLY(str(tmp_path / "m.ply"), PointCloud(mp, mn))
    assert writePLY(str(tmp_path / "s.ply"), PointCloud(w["scene"].points))
    exe = str(tmp_path / "facade")
    _compile(exe)
    out = subprocess.run([exe, str(tmp_path / "m.ply"), str(tmp_path / "s.ply"),
                          str(tmp_path / "x.model"), str(tmp_path),
                          os.path.join(ROOT, "tests", "golden", "rply_input.ply")],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "No trained model in matchScene" in out.stdout
    assert "featDistanceStepRel" in out.stdout
    assert "train model cost(ms):" in out.stdout
    assert "Read PLY failed: unable to open file" in out.stdout
    assert "mesh faces 1 samples 99" in out.stdout and "valid 99 box" in out.stdout
    lines = [l for l in out.stdout.splitlines() if l.startswith("score")]
    assert len(lines) >= 2
    D = 147.3
    for i, l in enumerate(lines):
        vals = [float(v) for v in l.split()[3:]]
        T = np.array(vals, np.float64).reshape(4, 4).T   # column-major
        k, ang, dt = synth.pose_error(T, w["scene"].gt_poses, D)
        assert ang < 1.0 and dt < 0.01 * D
        moved = PointCloud()
        assert readPLY(str(tmp_path / f"out{i}.ply"), moved)
        want = mp.astype(np.float64) @ T[:3, :3].T + T[:3, 3]
        assert np.abs(moved.point - want).max() < 1e-3
        assert np.allclose(moved.normal, mn @ T[:3, :3].T, atol=1e-5)
    n_sampled = int(out.stdout.split(" sampled ")[1].split()[0])
    ss = PointCloud()
    assert readPLY(str(tmp_path / "sampledScene.ply"), ss) and len(ss.point) == n_sampled
    kp = PointCloud()
    assert readPLY(str(tmp_path / "sampledKeypoint.ply"), kp) and len(kp.point) > 0
