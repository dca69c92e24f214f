"""oxr_trace_closest_device_noalloc (the closest-hit entry point on
device-resident rays that allocates nothing: its counter and spill live in
a fixed per-GPU device array) against oxr_trace_closest_device, the
entry point it is a variant of: the same hits bit for bit, on two calls in
a row, the second with fewer rays, for a shallow and a deeper tree."""
import ctypes as C

import numpy as np
import pytest

from oxr import _abi as A
from oxr import scenes

pytestmark = pytest.mark.gpu


def _rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = scene.bounds()
    c, ext = (lo + hi) / 2, (hi - lo) / 2 * 0.99
    o = (c + (rng.random((n, 3)) * 2 - 1) * ext).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.0, d, np.inf
    return r


class _Hip:
    """The HIP runtime liboxr.so itself uses (no second runtime in the
    process): device buffers for the rays and hits, which the caller owns."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes, fill=None):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        if fill is not None:
            assert self.h.hipMemset(p, C.c_int(fill), C.c_size_t(nbytes)) == 0
        return p

    def upload(self, arr):
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0  # H2D
        return p

    def download(self, p, arr):
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), p, C.c_size_t(arr.nbytes), 2) == 0  # D2H
        return arr

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("make", [lambda: scenes.material_sweep(segments=32, rings=16, floor_half=6),
                                  lambda: scenes.terrain(grid=192, n_emit=300, n_spheres=6000)],
                         ids=["sweep_small", "terrain_small"])
def test_noalloc_trace_equals_the_allocating_entry_point(make):
    import oxr
    from oxr._lib import check, lib
    if oxr.device_count() == 0:
        pytest.fail("no HIP device visible")
    L = lib()
    s = make()
    pt = oxr.PathTracer(0)
    pt.upload(s)
    hip = _Hip()
    for n, seed in ((1 << 17, 3), (3001, 4)):  # the second call is smaller
        r = _rays(s, n, seed)
        d_rays = hip.upload(np.ascontiguousarray(r.reshape(-1)))
        d_a = hip.alloc(n * 16, fill=0x5A)
        d_b = hip.alloc(n * 16, fill=0xA5)
        assert hip.h.hipDeviceSynchronize() == 0
        check(L.oxr_trace_closest_device(pt._h, n, d_rays, d_a, None), "trace_closest_device")
        check(L.oxr_trace_device_check(pt._h), "trace_device_check")
        check(L.oxr_trace_closest_device_noalloc(pt._h, n, d_rays, d_b, None), "trace_closest_device_noalloc")
        a = hip.download(d_a, np.zeros((n, 4), np.uint32))
        b = hip.download(d_b, np.zeros((n, 4), np.uint32))
        assert np.array_equal(a, b), (n, int((a != b).any(1).sum()))
        assert (a[:, 1] != A.OXR_HIT_MISS).mean() > 0.2
        # The host entry point traces the same rays to the same hits.
        host = pt.trace_closest((A.Ray * n).from_buffer_copy(r.tobytes()), n)
        assert np.array_equal(host, a)
    hip.close()
    pt.close()
