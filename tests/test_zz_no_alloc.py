"""ppf_voxel_sample_noalloc -- voxel sampling that allocates nothing (its
scratch is a fixed device array) -- against ppf_voxel_sample on the same
clouds: two calls in a row, the second on a smaller cloud (the scratch keeps
the first call's contents and must be cleared by the call itself), with and
without a subset, and the size limit.  Device buffers come from the HIP
runtime the library itself links (libamdhip64), through ctypes."""
import ctypes as C

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    from surface_match_amd import _native as N
    N.lib()  # loads libamdhip64 through the library
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class Dev:
    """device copies of host arrays, freed together"""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(int(nbytes), 4)) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _reference(pts, subset, step, box_min):
    from surface_match_amd import _native as N
    out = np.zeros(len(subset) if subset is not None else len(pts), np.int32)
    k = C.c_size_t()
    N.check(N.lib().ppf_voxel_sample(N.ptr(pts), len(pts), N.ptr(subset),
                                     0 if subset is None else len(subset), step, N.ptr(box_min),
                                     N.ptr(out), C.byref(k)))
    return out[:k.value]


def _noalloc(hip, pts, subset, step, box_min):
    from surface_match_amd import _native as N
    d = Dev(hip)
    try:
        x, y, z = (d.put(np.ascontiguousarray(pts[:, a])) for a in range(3))
        sub = d.put(subset) if subset is not None else None
        m = len(subset) if subset is not None else len(pts)
        out, cnt = d.alloc(4 * m), d.alloc(4)
        N.check(N.lib().ppf_voxel_sample_noalloc(x, y, z, len(pts), sub,
                                                 0 if subset is None else len(subset), step,
                                                 N.ptr(box_min), out, cnt, None))
        assert hip.hipDeviceSynchronize() == 0
        k = int(d.get(cnt, np.int32, 1)[0])
        return d.get(out, np.int32, k)
    finally:
        d.free()


def _cloud(n, seed):
    r = rng(seed)
    p = r.uniform(-100, 100, (n, 3)).astype(np.float32)
    p[: n // 4] = r.normal(scale=3.0, size=(n // 4, 3)).astype(np.float32)  # a dense blob
    return p


def test_noalloc_equals_voxel_sample_twice(hip):
    big, small = _cloud(300_000, 1), _cloud(50_000, 2)
    for pts, step in ((big, 1.5), (small, 4.0)):
        box = pts.min(0).astype(np.float32)
        ref = _reference(pts, None, step, box)
        got = _noalloc(hip, pts, None, step, box)
        assert len(ref) > 0 and np.array_equal(got, ref), (len(got), len(ref))
    # a subset of the small cloud after the big one
    sub = np.sort(rng(3).choice(len(small), 20_000, replace=False)).astype(np.int32)
    box = small.min(0).astype(np.float32)
    assert np.array_equal(_noalloc(hip, small, sub, 2.0, box), _reference(small, sub, 2.0, box))


def test_noalloc_limit_and_empty(hip):
    from surface_match_amd import _native as N
    pts = _cloud(1000, 4)
    box = pts.min(0).astype(np.float32)
    d = Dev(hip)
    try:
        x, y, z = (d.put(np.ascontiguousarray(pts[:, a])) for a in range(3))
        out, cnt = d.alloc(4 * 1000), d.alloc(4)
        # more points than the fixed scratch holds: rejected before any launch
        rc = N.lib().ppf_voxel_sample_noalloc(x, y, z, (1 << 20) + 1, None, 0, 1.0,
                                              N.ptr(box), out, cnt, None)
        assert rc == 1 and b"at most 1048576" in N.lib().ppf_last_error()
        # no points: count 0
        assert N.lib().ppf_voxel_sample_noalloc(x, y, z, 0, None, 0, 1.0, N.ptr(box), out, cnt,
                                                None) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert int(d.get(cnt, np.int32, 1)[0]) == 0
    finally:
        d.free()
