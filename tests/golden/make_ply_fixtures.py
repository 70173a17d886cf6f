"""Writes small PLY files in every storage format and the reference's read of
each (readPLY over rply, filePLY.cpp:199-257, through the harness that
oracle/Makefile builds from the reference tree into oracle/_ref) as
tests/golden/ply/expected.npz, so the GPU box, which has no reference tree,
can check the library's PLY upload against the reference's result.

    make -C oracle ref && python tests/golden/make_ply_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

OUT = os.path.join(HERE, "ply")


def main():
    import ctypes as C
    from conftest import rng
    from test_ply import REF, cloud_rows, ref_read, write_ply
    R = C.CDLL(REF)
    R.plyref_read.argtypes = [C.c_char_p] + [C.POINTER(C.c_long)] * 3 + \
        [C.POINTER(C.POINTER(C.c_float))] * 2 + [C.POINTER(C.POINTER(C.c_int))]
    R.plyref_free.argtypes = [C.c_void_p]
    os.makedirs(OUT, exist_ok=True)
    files = {}
    r = rng(21)
    rows = cloud_rows(200, 22)
    vprops = [("float", "x"), ("float", "y"), ("float", "z"), ("float", "nx"), ("float", "ny"),
              ("float", "nz")]
    faces = [list(r.choice(200, 3, replace=False)) for _ in range(30)]
    for fmt, tag in (("ascii", "ascii"), ("binary_little_endian", "le"),
                     ("binary_big_endian", "be")):
        name = f"normals_{tag}.ply"
        write_ply(os.path.join(OUT, name), fmt, vprops, rows, faces)
        files[name] = True
    # double coordinates, no normals, an interleaved property, vertex_index
    rows2 = [[x * 1.0000001, int(r.integers(0, 256)), y, z] for x, y, z, *_ in cloud_rows(150, 23)]
    write_ply(os.path.join(OUT, "double_plain_le.ply"), "binary_little_endian",
              [("double", "x"), ("uchar", "red"), ("double", "y"), ("double", "z")], rows2,
              [[0, 1, 2], [3, 4, 5]], face_prop=("uint8", "uint32", "vertex_index"))
    files["double_plain_le.ply"] = True
    expected = {}
    for name in sorted(files):
        pts, nrm, tri = ref_read(R, os.path.join(OUT, name))
        stem = name[:-4]
        expected[stem + "/points"] = pts
        expected[stem + "/normals"] = nrm
        expected[stem + "/faces"] = tri
    np.savez(os.path.join(OUT, "expected.npz"), **expected)
    print(sorted(files))


if __name__ == "__main__":
    main()
