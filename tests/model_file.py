"""A reader and writer of the reference's model file, written from its
serializer alone (serialize.cpp:5-186 and Detector::save / load,
ppf.cpp:512-547) -- independent of the library's own save and load, so the
tests can check the library's files field by field and hand it files laid
out as the reference may write them.

Layout (packed, little endian):
    int32 MAGIC 0x7F27F, int32 VERSION 100         (ppf.cpp:21-22, 519-520)
    float samplingDistanceRel                       (ppf.cpp:521)
    TrainParam: float, int32, float, int32, bool(1 byte)   (serialize.cpp:180-186)
    PointCloud sampledModel, reSampledModel:        (serialize.cpp:72-84)
        uint32 n, n x float[3] points; uint32 m, m x float[3] normals;
        float[3] box.min, float[3] box.max; float[3] viewPoint
    hash table: uint32 size, size x {uint32 key; uint32 n, n x uint32 refInd;
        uint32 m, m x float alphaAngle}             (serialize.cpp:96-170)
"""
import struct

import numpy as np

MAGIC, VERSION = 0x7F27F, 100


class _Reader:
    def __init__(self, data):
        self.b, self.o = data, 0

    def take(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.o)
        self.o += struct.calcsize("<" + fmt)
        return v if len(v) > 1 else v[0]

    def array(self, dtype, count):
        a = np.frombuffer(self.b, dtype, count, self.o).copy()
        self.o += a.nbytes
        return a


def _read_cloud(r):
    n = r.take("I")
    p = r.array("<f4", 3 * n).reshape(n, 3)
    m = r.take("I")
    nrm = r.array("<f4", 3 * m).reshape(m, 3)
    box_min = r.array("<f4", 3)
    box_max = r.array("<f4", 3)
    vp = r.array("<f4", 3)
    return dict(point=p, normal=nrm, box_min=box_min, box_max=box_max, view_point=vp)


def read_model(path):
    with open(path, "rb") as f:
        r = _Reader(f.read())
    out = {"magic": r.take("i"), "version": r.take("i"), "samplingDistanceRel": r.take("f")}
    out["param"] = dict(zip(("featDistanceStepRel", "featAngleResolution",
                             "poseRefRelSamplingDistance", "knnNormal", "smoothNormal"),
                            r.take("fifi?")))
    out["sampled"] = _read_cloud(r)
    out["resampled"] = _read_cloud(r)
    table = []
    for _ in range(r.take("I")):
        key = r.take("I")
        refs = r.array("<u4", r.take("I"))
        alphas = r.array("<f4", r.take("I"))
        table.append((key, refs, alphas))
    out["table"] = table
    out["trailing_bytes"] = len(r.b) - r.o
    return out


def _write_cloud(f, c):
    p = np.ascontiguousarray(c["point"], "<f4").reshape(-1, 3)
    n = np.ascontiguousarray(c["normal"], "<f4").reshape(-1, 3)
    f.write(struct.pack("<I", len(p)) + p.tobytes())
    f.write(struct.pack("<I", len(n)) + n.tobytes())
    for k in ("box_min", "box_max", "view_point"):
        f.write(np.ascontiguousarray(c[k], "<f4").reshape(3).tobytes())


def write_model(path, m):
    """m as read_model returns it; the table is written in the given order
    (the reference writes its hash map's iteration order)."""
    with open(path, "wb") as f:
        f.write(struct.pack("<iif", m.get("magic", MAGIC), m.get("version", VERSION),
                            m["samplingDistanceRel"]))
        p = m["param"]
        f.write(struct.pack("<fifi?", p["featDistanceStepRel"], p["featAngleResolution"],
                            p["poseRefRelSamplingDistance"], p["knnNormal"],
                            bool(p["smoothNormal"])))
        _write_cloud(f, m["sampled"])
        _write_cloud(f, m["resampled"])
        f.write(struct.pack("<I", len(m["table"])))
        for key, refs, alphas in m["table"]:
            refs = np.ascontiguousarray(refs, "<u4")
            alphas = np.ascontiguousarray(alphas, "<f4")
            f.write(struct.pack("<II", key, len(refs)) + refs.tobytes())
            f.write(struct.pack("<I", len(alphas)) + alphas.tobytes())
