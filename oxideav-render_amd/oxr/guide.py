"""Learned path guide over liboxr.so (include/oxr/guide.h).

A hash-grid + MLP (BF16 MFMA, FP32 accumulate) proposal of incident
directions, trained online on the tracer's own radiance samples and
combined with BSDF sampling by one-sample MIS. oxideav-render has no path
guiding; with no guide attached the tracer is the reference estimator.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._lib import check, lib

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)


def default_config(**overrides):
    """oxr_guide_default_config, then field overrides (aabb_min/aabb_max
    as 3-tuples)."""
    c = A.GuideConfig()
    check(lib().oxr_guide_default_config(C.byref(c)), "guide_default_config")
    for k, v in overrides.items():
        if k in ("aabb_min", "aabb_max"):
            setattr(c, k, A.f3(*v))
        else:
            setattr(c, k, v)
    return c


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if n is None else a.reshape(n)


class Guide:
    def __init__(self, device=0, config=None, **overrides):
        self._L = lib()
        self.config = config if config is not None else default_config(**overrides)
        h = C.c_void_p()
        check(self._L.oxr_guide_create(device, C.byref(self.config), C.byref(h)), "oxr_guide_create")
        self._h = h
        ng, nm = C.c_uint64(0), C.c_uint64(0)
        check(self._L.oxr_guide_num_params(h, C.byref(ng), C.byref(nm)), "num_params")
        self.n_grid, self.n_mlp = ng.value, nm.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            self._L.oxr_guide_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def params(self):
        grid = np.zeros(self.n_grid, np.float32)
        mlp = np.zeros(self.n_mlp, np.float32)
        check(self._L.oxr_guide_get_params(self._h, grid.ctypes.data_as(fp), mlp.ctypes.data_as(fp)), "get_params")
        return grid, mlp

    def set_params(self, grid=None, mlp=None):
        g = _f32(grid) if grid is not None else None
        m = _f32(mlp) if mlp is not None else None
        if g is not None and g.size != self.n_grid or m is not None and m.size != self.n_mlp:
            raise ValueError("parameter sizes do not match the guide")
        check(self._L.oxr_guide_set_params(self._h, g.ctypes.data_as(fp) if g is not None else fp(),
                                           m.ctypes.data_as(fp) if m is not None else fp()), "set_params")

    def infer(self, pos):
        pos = _f32(pos).reshape(-1, 3)
        n = len(pos)
        raw = np.zeros((n, A.GUIDE_OUTPUTS), np.float32)
        lobes = np.zeros((n, A.GUIDE_LOBES, 5), np.float32)
        check(self._L.oxr_guide_infer(self._h, n, pos.ctypes.data_as(fp), raw.ctypes.data_as(fp),
                                      lobes.ctypes.data_as(fp)), "guide_infer")
        return raw, lobes

    def pdf(self, pos, dirs):
        pos, dirs = _f32(pos).reshape(-1, 3), _f32(dirs).reshape(-1, 3)
        out = np.zeros(len(pos), np.float32)
        check(self._L.oxr_guide_pdf(self._h, len(pos), pos.ctypes.data_as(fp), dirs.ctypes.data_as(fp),
                                    out.ctypes.data_as(fp)), "guide_pdf")
        return out

    def sample(self, pos, u):
        pos, u = _f32(pos).reshape(-1, 3), _f32(u).reshape(-1, 3)
        d = np.zeros((len(pos), 3), np.float32)
        p = np.zeros(len(pos), np.float32)
        check(self._L.oxr_guide_sample(self._h, len(pos), pos.ctypes.data_as(fp), u.ctypes.data_as(fp),
                                       d.ctypes.data_as(fp), p.ctypes.data_as(fp)), "guide_sample")
        return d, p

    def gradient(self, pos, dirs, weight):
        pos, dirs, w = _f32(pos).reshape(-1, 3), _f32(dirs).reshape(-1, 3), _f32(weight).ravel()
        gg = np.zeros(self.n_grid, np.float32)
        gm = np.zeros(self.n_mlp, np.float32)
        loss = C.c_double(0.0)
        check(self._L.oxr_guide_gradient(self._h, len(pos), pos.ctypes.data_as(fp), dirs.ctypes.data_as(fp),
                                         w.ctypes.data_as(fp), gg.ctypes.data_as(fp), gm.ctypes.data_as(fp),
                                         C.byref(loss)), "guide_gradient")
        return loss.value, gg, gm

    def train(self, pos, dirs, weight, steps=1):
        pos, dirs, w = _f32(pos).reshape(-1, 3), _f32(dirs).reshape(-1, 3), _f32(weight).ravel()
        loss = C.c_double(0.0)
        check(self._L.oxr_guide_train(self._h, len(pos), pos.ctypes.data_as(fp), dirs.ctypes.data_as(fp),
                                      w.ctypes.data_as(fp), steps, C.byref(loss)), "guide_train")
        return loss.value

    def collected(self, max_samples=1 << 20):
        """Training samples of the last online pass: (pos (n,3), weight (n,),
        dir (n,3), proposal pdf (n,))."""
        n = C.c_uint32(0)
        check(self._L.oxr_guide_collected(self._h, 0, None, C.byref(n)), "guide_collected")
        m = min(n.value, max_samples)
        out = np.zeros((max(m, 1), 8), np.float32)
        check(self._L.oxr_guide_collected(self._h, m, out.ctypes.data_as(fp), C.byref(n)), "guide_collected")
        out = out[:m]
        return out[:, 0:3], out[:, 3], out[:, 4:7], out[:, 7]

    def stats(self):
        s = A.GuideStats()
        check(self._L.oxr_guide_stats_get(self._h, C.byref(s)), "guide_stats")
        return {f: getattr(s, f) for f, _ in A.GuideStats._fields_ if f != "pad"}
