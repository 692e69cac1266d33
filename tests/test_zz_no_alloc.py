"""mcgpu_broad_phase_noalloc (the broad phase over a static scratch array, no
device allocation) against mcgpu_broad_phase on the same face boxes, through
the C ABI: two calls in a row, the second on a smaller input, must give the
pair sets of the allocating entry point exactly."""
import ctypes as C

import numpy as np
import pytest

import meshes


def face_boxes(mesh, eps=0.0):
    V, F, S = mesh
    assert S is None
    P = V[F.reshape(-1, 3)]
    lo, hi = P.min(axis=1), P.max(axis=1)
    return np.ascontiguousarray(np.hstack([lo - eps, hi + eps]))


def box_pair(pair):
    src, cut = pair
    d = cut[0].max(axis=0) - cut[0].min(axis=0)
    eps = float(np.sqrt((d * d).sum())) * 1e-4
    return face_boxes(src), face_boxes(cut, eps)


def test_sizes_beyond_the_maximum_are_refused_before_device_work():
    from mcut_amd import frontend
    big = frontend.NOALLOC_MAX_FACES + 1
    # non-null stand-in pointers: the size check comes before any use of them
    r, n = frontend.broad_phase_noalloc(0, 8, big, 8, 10, 8, 1, None)
    assert r == -1 and n == 0
    r, n = frontend.broad_phase_noalloc(0, 8, 10, 8, big, 8, 1, None)
    assert r == -1
    r, n = frontend.broad_phase_noalloc(0, None, 10, 8, 10, 8, 1, None)
    assert r == -1


class Hip:
    """The HIP runtime libmcgpu links, through ctypes: the test (the caller)
    owns its input and output buffers."""

    def __init__(self):
        from mcut_amd import frontend
        frontend.lib()  # loads libamdhip64 first; the same runtime is reused here
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, a):
        assert self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # device to host
        return a

    def free(self, p):
        assert self.L.hipFree(p) == 0


@pytest.mark.gpu
def test_noalloc_matches_the_allocating_broad_phase():
    from mcut_amd import frontend
    hip = Hip()
    ctx = frontend.Context(0)
    cases = [("icosphere5", meshes.icosphere_pair(5, amp=0.02)), ("config1_rotated", meshes.config1_rotated())]
    cap = 1 << 20
    sentinel = np.full(cap, np.uint64(0xffffffffffffffff), np.uint64)
    out = hip.upload(sentinel)
    bufs = []
    try:
        for name, pair in cases:  # the second call is the smaller input
            print(f"noalloc case {name}", flush=True)
            a, b = box_pair(pair)
            want = ctx.broad_phase(a, b)
            da, db = hip.upload(a), hip.upload(b)
            bufs += [da, db]
            assert hip.L.hipMemcpy(out, sentinel.ctypes.data, sentinel.nbytes, 1) == 0
            r, n = frontend.broad_phase_noalloc(0, da, a.shape[0], db, b.shape[0], out, cap, None)
            assert r == 0, r
            assert n == want.shape[0] and n > 0, (n, want.shape)
            res = hip.download(out, np.zeros(cap, np.uint64))
            got = res[:n]
            got = np.stack([(got >> np.uint64(32)).astype(np.uint32), (got & np.uint64(0xffffffff)).astype(np.uint32)],
                           axis=1)
            got = got[np.lexsort((got[:, 1], got[:, 0]))]
            assert np.array_equal(got, want), name
            assert np.all(res[n:] == np.uint64(0xffffffffffffffff)), "written past the pair count"
            # a capacity below the count is refused with the count
            r, m = frontend.broad_phase_noalloc(0, da, a.shape[0], db, b.shape[0], out, n - 1, None)
            assert r == -3 and m == n
    finally:
        ctx.close()
        for p in bufs + [out]:
            hip.free(p)
