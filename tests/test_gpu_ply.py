"""PLY ingest on the GPU box against the reference's read (readPLY over rply,
filePLY.cpp:199-257): each fixture file (ascii, binary little and big
endian, float and double coordinates, with and without normals, an
interleaved property; tests/golden/make_ply_fixtures.py) is uploaded as a
resident scene with ppf_scene_upload_ply, and the scene's points and
normals come back from HBM bit for bit equal to what the reference read;
the triangles from ppf_ply_read equal the reference's too."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    exp = np.load(os.path.join(GOLD, "ply", "expected.npz"))
    stems = sorted({k.split("/")[0] for k in exp.files})
    out = [(os.path.join(GOLD, "ply", s + ".ply"), exp[s + "/points"], exp[s + "/normals"],
            exp[s + "/faces"]) for s in stems]
    # rply's own sample file (ASCII, CRLF, leading zeros): known answer
    out.append((os.path.join(GOLD, "rply_input.ply"),
                np.array([[-1, 0, 0], [0, 1, 0], [1, 0, 0]], np.float32),
                np.zeros((0, 3), np.float32), np.array([[1, 0, 2]], np.int32)))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: os.path.basename(c[0]))
def test_ply_upload_equals_reference_read(case):
    from surface_match_amd import Detector, PointCloud, ResidentScene, readPLY
    path, pts, nrm, faces = case
    rs = ResidentScene(Detector(), ply=path)
    gp, gn, valid, box = rs.download()
    assert np.array_equal(gp.view(np.uint32), pts.astype(np.float32).view(np.uint32))
    if len(nrm):
        assert gn is not None and np.array_equal(gn.view(np.uint32),
                                                 nrm.astype(np.float32).view(np.uint32))
    else:
        assert gn is None
    assert np.array_equal(valid, np.arange(len(pts)))
    m = PointCloud()
    assert readPLY(path, m)
    assert np.array_equal(m.face, faces)
