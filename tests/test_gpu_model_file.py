"""Model files (Detector::save / load, ppf.cpp:512-547, serialize.cpp:5-186)
checked with a reader and writer written from the reference's serializer
alone (tests/model_file.py), not with the library's own code:

* every field of a saved file against the trained model;
* a file whose buckets come in descending key order, as the reference's
  hash map may write them, loads to the same table and the same matches;
* a key stored twice keeps its first bucket (flat_hash_map::emplace);
* corrupt files are rejected with std::runtime_error / range_error and
  leave the detector without a model.
"""
import math

import numpy as np
import pytest

import model_file as MF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trained():
    from surface_match_amd import Detector, PointCloud, synth
    w = synth.workload(1, scale=0.5)
    det = Detector()
    det.trainModel(PointCloud(*w["model"]), 0.04)
    cloud = PointCloud(w["scene"].points)
    cloud.viewPoint = np.array(synth.VIEWPOINT, np.float32)
    return det, cloud


@pytest.fixture
def saved(trained, tmp_path):
    det, _ = trained
    f = tmp_path / "m.model"
    det.save(str(f))
    return f


def test_saved_file_fields(trained, saved):
    det, _ = trained
    m = MF.read_model(saved)
    assert m["magic"] == MF.MAGIC and m["version"] == MF.VERSION and m["trailing_bytes"] == 0
    assert m["samplingDistanceRel"] == np.float32(0.04)
    assert m["param"] == {"featDistanceStepRel": np.float32(0.04), "featAngleResolution": 30,
                          "poseRefRelSamplingDistance": np.float32(0.01), "knnNormal": 10,
                          "smoothNormal": True}
    sp, sn, rp, rn = det.sampled_model()
    info = det.model_info()
    for cloud, P, Nn in ((m["sampled"], sp, sn), (m["resampled"], rp, rn)):
        assert np.array_equal(cloud["point"].view(np.uint32), P.view(np.uint32))
        assert np.array_equal(cloud["normal"].view(np.uint32), Nn.view(np.uint32))
        assert np.array_equal(cloud["box_min"], np.array(info.boxMin, np.float32))
        assert np.array_equal(cloud["box_max"], np.array(info.boxMax, np.float32))
        assert np.isnan(cloud["view_point"]).all()  # the model had no view point
    keys, offs, refs, alphas = det.table()
    assert [k for k, _, _ in m["table"]] == keys.tolist()
    for b, (key, r, a) in enumerate(m["table"]):
        assert np.array_equal(r, refs[offs[b]:offs[b + 1]])
        assert np.array_equal(a.view(np.uint32), alphas[offs[b]:offs[b + 1]].view(np.uint32))


def test_buckets_in_descending_key_order(trained, saved, tmp_path):
    from surface_match_amd import Detector, MatchParam
    det, cloud = trained
    m = MF.read_model(saved)
    m["table"] = sorted(m["table"], key=lambda t: -t[0])
    f = tmp_path / "desc.model"
    MF.write_model(f, m)
    d2 = Detector()
    d2.load(str(f))
    for x, y in zip(det.table(), d2.table()):
        assert np.array_equal(x, y)
    param = MatchParam(numMatches=3)
    P1, S1 = det.matchScene(cloud, 0.04, 0.1, 0.2, param)
    P2, S2 = d2.matchScene(cloud, 0.04, 0.1, 0.2, param)
    assert len(S1) > 0 and np.array_equal(P1, P2) and np.array_equal(S1, S2)


def test_repeated_key_keeps_first_bucket(trained, saved, tmp_path):
    from surface_match_amd import Detector
    det, _ = trained
    m = MF.read_model(saved)
    key, refs, alphas = m["table"][5]
    m["table"].append((key, refs[:1], alphas[:1]))  # a later, different bucket for the key
    f = tmp_path / "dup.model"
    MF.write_model(f, m)
    d2 = Detector()
    d2.load(str(f))
    for x, y in zip(det.table(), d2.table()):
        assert np.array_equal(x, y)


def _corrupt_cases(m):
    import copy
    out = []
    c = copy.deepcopy(m)
    c["table"][3] = (c["table"][3][0], c["table"][3][1].copy(), c["table"][3][2])
    c["table"][3][1][0] = len(m["sampled"]["point"])          # reference past the model
    out.append(("ref out of range", c, RuntimeError, "corrupt hash table"))
    c = copy.deepcopy(m)
    c["sampled"]["normal"] = c["sampled"]["normal"][:-1]      # one normal short
    out.append(("normals short", c, RuntimeError, "corrupt point cloud"))
    c = copy.deepcopy(m)
    c["param"]["featAngleResolution"] = 0
    out.append(("angle resolution 0", c, ValueError, "featAngleResolution"))
    c = copy.deepcopy(m)
    k, r, a = c["table"][0]
    a = a.copy()
    a[0] = np.nan
    c["table"][0] = (k, r, a)
    out.append(("NaN alpha", c, RuntimeError, "corrupt hash table"))
    c = copy.deepcopy(m)
    c["sampled"]["point"] = c["sampled"]["point"][:0]
    c["sampled"]["normal"] = c["sampled"]["normal"][:0]
    out.append(("empty model", c, RuntimeError, "corrupt point cloud"))
    return out


def test_corrupt_files_are_rejected(saved, tmp_path):
    import struct
    from surface_match_amd import Detector, PointCloud
    m = MF.read_model(saved)
    cases = _corrupt_cases(m)
    data = saved.read_bytes()
    raw = [("truncated", data[: len(data) // 2], RuntimeError, "truncated model file"),
           # a point count near 2^32 right after the parameters
           ("huge count", data[:29] + struct.pack("<I", 0xFFFFFFF0) + data[33:], RuntimeError,
            "truncated model file"),
           ("bad magic", b"\0\0\0\0" + data[4:], RuntimeError, "unsupported file format")]
    for name, c, exc, msg in cases:
        f = tmp_path / (name.replace(" ", "_") + ".model")
        MF.write_model(f, c)
        raw.append((name, f.read_bytes(), exc, msg))
    for name, blob, exc, msg in raw:
        f = tmp_path / "case.model"
        f.write_bytes(blob)
        d = Detector()
        with pytest.raises(exc, match=msg):
            d.load(str(f))
        # the failed load leaves no model behind
        with pytest.raises(RuntimeError, match="No trained model"):
            d.matchScene(PointCloud(np.zeros((10, 3), np.float32)))
        print(f"{name}: rejected")
