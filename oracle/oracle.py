"""ctypes front of the oracle (liboxr_cpu.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and the bench's
cpu_baseline leg use this module, as the checker and the host-CPU
baseline. Nothing in the product imports it.
"""
import ctypes as C
import os

import numpy as np

from oxr import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

fp = C.POINTER(C.c_float)
dp = C.POINTER(C.c_double)
up = C.POINTER(C.c_uint32)
bp = C.POINTER(C.c_uint8)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboxr_cpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.oxrcpu_scene_create.restype = C.c_void_p
        L.oxrcpu_scene_create.argtypes = [C.POINTER(A.SceneDesc)]
        L.oxrcpu_scene_destroy.argtypes = [C.c_void_p]
        L.oxrcpu_num_lights.restype = C.c_uint32
        L.oxrcpu_num_lights.argtypes = [C.c_void_p]
        L.oxrcpu_emissive_table.argtypes = [C.c_void_p, up, fp, fp, up]
        L.oxrcpu_render.argtypes = [C.c_void_p, C.POINTER(A.Camera), C.POINTER(A.PtOptions), C.c_uint32,
                                    C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, dp, dp]
        L.oxrcpu_render_rect.argtypes = [C.c_void_p, C.POINTER(A.Camera), C.POINTER(A.PtOptions), C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, dp, dp]
        L.oxrcpu_sample_one.argtypes = [C.c_void_p, C.POINTER(A.Camera), C.POINTER(A.PtOptions), C.c_uint32,
                                        C.c_uint32, C.c_uint32, fp]
        L.oxrcpu_trace_closest.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(A.Ray), C.c_uint32, C.c_int,
                                           C.POINTER(A.Hit)]
        L.oxrcpu_trace_any.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(A.Ray), C.c_int, bp]
        L.oxrcpu_sobol_owen_4d.argtypes = [C.c_uint32, up, up, fp]
        L.oxrcpu_sobol.restype = C.c_uint32
        L.oxrcpu_sobol.argtypes = [C.c_uint32, C.c_int]
        L.oxrcpu_pcg_hash.restype = C.c_uint32
        L.oxrcpu_pcg_hash.argtypes = [C.c_uint32]
        L.oxrcpu_sobol_matrices.argtypes = [up]
        L.oxrcpu_sheen_lut.argtypes = [fp]
        L.oxrcpu_bsdf_eval.argtypes = [C.c_uint32, C.POINTER(A.MaterialSample), fp, fp, bp, fp, fp, fp]
        L.oxrcpu_bsdf_sample.argtypes = [C.c_uint32, C.POINTER(A.MaterialSample), fp, fp, bp, fp, fp, fp, fp, bp]
        L.oxrcpu_bsdf_probs.argtypes = [C.POINTER(A.MaterialSample), fp, fp, C.c_uint8, fp]
        L.oxrcpu_triangle_solid_angle.restype = C.c_float
        L.oxrcpu_triangle_solid_angle.argtypes = [fp, fp, fp, fp]
        L.oxrcpu_sample_spherical_triangle.restype = C.c_int
        L.oxrcpu_sample_spherical_triangle.argtypes = [fp, fp, fp, fp, C.c_float, C.c_float, fp]
        L.oxrcpu_iridescent_fresnel.argtypes = [C.c_float, C.c_float, fp, C.c_float, C.c_float, fp]
        L.oxrcpu_offset_ray_origin.argtypes = [fp, fp, fp]
        L.oxrcpu_set_environment.argtypes = [C.c_void_p, fp, C.c_uint32, C.c_uint32, C.c_float]
        L.oxrcpu_env_eval.argtypes = [C.c_void_p, C.c_uint32, fp, fp, fp]
        L.oxrcpu_env_sample.argtypes = [C.c_void_p, C.c_uint32, fp, fp, fp, bp]
        L.oxrcpu_env_tables.argtypes = [C.c_void_p, fp, fp, fp, bp]
        gc = C.POINTER(A.GuideConfig)
        L.oxrcpu_guide_resolutions.argtypes = [gc, up, bp]
        L.oxrcpu_bf16_round.restype = C.c_float
        L.oxrcpu_bf16_round.argtypes = [C.c_float]
        L.oxrcpu_guide_corners.argtypes = [gc, C.c_uint32, fp, up, fp]
        L.oxrcpu_guide_forward.argtypes = [gc, fp, fp, C.c_uint32, fp, C.c_int, fp, fp]
        L.oxrcpu_guide_lobes.argtypes = [C.c_uint32, fp, fp]
        L.oxrcpu_guide_pdf.argtypes = [C.c_uint32, fp, fp, dp]
        L.oxrcpu_guide_sample.argtypes = [C.c_uint32, fp, fp, fp]
        L.oxrcpu_guide_loss_grad.restype = C.c_double
        L.oxrcpu_guide_loss_grad.argtypes = [gc, fp, fp, C.c_uint32, fp, fp, fp, C.c_int, dp, dp]
        L.oxrcpu_guide_adam.argtypes = [C.c_uint64, fp, fp, fp, fp, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.c_int]
        L.oxrcpu_barycentric_of.restype = C.c_int
        L.oxrcpu_barycentric_of.argtypes = [fp, fp, fp, fp, fp]
        L.oxrcpu_surface.argtypes = [C.c_void_p, C.c_uint32, fp, fp]
        L.oxrcpu_resolve_hdr.argtypes = [dp, C.c_uint32, C.c_uint32, C.POINTER(A.Display), fp]
        L.oxrcpu_resolve_rgba8.argtypes = [dp, C.c_uint32, C.c_uint32, C.POINTER(A.Display), bp]
        L.oxrcpu_tone_map.argtypes = [C.c_uint32, C.c_uint32, fp, fp]
        L.oxrcpu_linear_to_srgb_byte.argtypes = [C.c_uint32, fp, bp]
        L.oxrcpu_srgb_to_linear.argtypes = [C.c_uint32, fp, fp]
        L.oxrcpu_record_rays.restype = C.c_uint32
        L.oxrcpu_record_rays.argtypes = [C.c_void_p, C.POINTER(A.Camera), C.POINTER(A.PtOptions), C.c_uint32,
                                         C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, fp]
        L.oxrcpu_bvh4_count.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, fp, up]
        L.oxrcpu_bvh4_stack_depth.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, fp, up]
        L.oxrcpu_bvh4_count_any.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, fp, up]
        L.oxrcpu_record_shadow_rays.restype = C.c_uint32
        L.oxrcpu_record_shadow_rays.argtypes = [C.c_void_p, C.POINTER(A.Camera), C.POINTER(A.PtOptions), C.c_uint32,
                                                C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, fp]
        L.oxrcpu_tex_sample.restype = C.c_int
        L.oxrcpu_tex_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32, fp, fp, fp]
        L.oxrcpu_tex_grad_lod.restype = C.c_int
        L.oxrcpu_tex_grad_lod.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, fp, fp]
        L.oxrcpu_tex_level.restype = C.c_int
        L.oxrcpu_tex_level.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_uint32, up, fp]
        L.oxrcpu_material_eval.restype = C.c_int
        L.oxrcpu_material_eval.argtypes = [C.c_void_p, C.c_uint32, up, fp, fp, fp, bp, fp]
        _LIB = L
    return _LIB


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(fp)


def options(width, height, spp=1, max_bounces=8, rr_start=3, clamp=0.0, seed=0, strategy=0, ambient=0.0):
    o = A.PtOptions()
    o.width, o.height = width, height
    o.samples_per_pixel = spp
    o.max_bounces = max_bounces
    o.rr_start = rr_start
    o.clamp = clamp
    o.seed = seed
    o.strategy = strategy
    o.ambient = ambient
    return o


class OracleScene:
    """The oracle's prepared scene (soup + its own binned-SAH BVH)."""

    def __init__(self, scene):
        self.scene = scene
        self._desc = scene.desc()
        self.h = lib().oxrcpu_scene_create(C.byref(self._desc))
        if not self.h:
            raise RuntimeError("oracle scene creation failed")

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.oxrcpu_scene_destroy(self.h)
            self.h = None

    def render(self, width, height, first, n, opts=None, y0=0, y1=None, threads=0, sum_sq=False, cam=None):
        """Accumulate samples [first, first+n) into fresh f64 buffers:
        returns accum (H', W, 4) [and sum_sq (H', W, 3)]."""
        y1 = height if y1 is None else y1
        if opts is None:
            opts = options(width, height, ambient=getattr(self.scene, "ambient", 0.0))
        cam = cam or self.scene.camera.resolve(width, height)
        acc = np.zeros((y1 - y0, width, 4), dtype=np.float64)
        sq = np.zeros((y1 - y0, width, 3), dtype=np.float64) if sum_sq else None
        rc = lib().oxrcpu_render(self.h, C.byref(cam), C.byref(opts), first, n, y0, y1, threads,
                                 acc.ctypes.data_as(dp), sq.ctypes.data_as(dp) if sq is not None else dp())
        if rc != 0:
            raise RuntimeError("oracle render failed")
        return (acc, sq) if sum_sq else acc

    def render_rect(self, width, height, first, n, x0, x1, y0, y1, opts=None, threads=0, cam=None):
        """Samples [first, first+n) of the pixels in columns [x0, x1) and
        rows [y0, y1) of a width x height frame: (accum (h, w, 4),
        sum_sq (h, w, 3)), f64 sums in sample order."""
        if opts is None:
            opts = options(width, height, ambient=getattr(self.scene, "ambient", 0.0))
        cam = cam or self.scene.camera.resolve(width, height)
        acc = np.zeros((y1 - y0, x1 - x0, 4), dtype=np.float64)
        sq = np.zeros((y1 - y0, x1 - x0, 3), dtype=np.float64)
        rc = lib().oxrcpu_render_rect(self.h, C.byref(cam), C.byref(opts), first, n, x0, x1, y0, y1, threads,
                                      acc.ctypes.data_as(dp), sq.ctypes.data_as(dp))
        if rc != 0:
            raise RuntimeError("oracle render_rect failed")
        return acc, sq

    def sample_one(self, width, height, x, y, index, opts=None, cam=None):
        if opts is None:
            opts = options(width, height, ambient=getattr(self.scene, "ambient", 0.0))
        cam = cam or self.scene.camera.resolve(width, height)
        rgb = np.zeros(3, np.float32)
        cov = lib().oxrcpu_sample_one(self.h, C.byref(cam), C.byref(opts), x, y, index, rgb.ctypes.data_as(fp))
        return (rgb if cov else None)

    def trace_closest(self, rays, flags=0, brute=False):
        n = len(rays)
        hits = (A.Hit * max(n, 1))()
        lib().oxrcpu_trace_closest(self.h, n, rays, flags, 1 if brute else 0, hits)
        return hits_to_numpy(hits, n)

    def trace_any(self, rays, brute=False):
        n = len(rays)
        occ = np.zeros(max(n, 1), np.uint8)
        lib().oxrcpu_trace_any(self.h, n, rays, 1 if brute else 0, occ.ctypes.data_as(bp))
        return occ[:n]

    def set_environment(self, rgba, intensity=1.0):
        """PathTracer::set_environment: rgba (H, W, 4) scene-linear, row 0 =
        +Y; None removes it."""
        if rgba is None:
            lib().oxrcpu_set_environment(self.h, fp(), 0, 0, 0.0)
            self._env = None
            return
        img = np.ascontiguousarray(rgba, dtype=np.float32)
        self._env = img
        lib().oxrcpu_set_environment(self.h, img.ctypes.data_as(fp), img.shape[1], img.shape[0], float(intensity))

    def env_eval(self, dirs):
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        pdf = np.zeros(len(d), np.float32)
        le = np.zeros((len(d), 3), np.float32)
        lib().oxrcpu_env_eval(self.h, len(d), d.ctypes.data_as(fp), pdf.ctypes.data_as(fp), le.ctypes.data_as(fp))
        return pdf, le

    def env_sample(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        d = np.zeros((len(u), 3), np.float32)
        pdf = np.zeros(len(u), np.float32)
        ok = np.zeros(len(u), np.uint8)
        lib().oxrcpu_env_sample(self.h, len(u), u.ctypes.data_as(fp), d.ctypes.data_as(fp), pdf.ctypes.data_as(fp),
                                ok.ctypes.data_as(bp))
        return d, pdf, ok.astype(bool)

    def env_tables(self):
        h, w = self._env.shape[:2]
        mg = np.zeros(h + 1, np.float32)
        cd = np.zeros(h * (w + 1), np.float32)
        ms = np.zeros(h * w, np.float32)
        un = np.zeros(1, np.uint8)
        lib().oxrcpu_env_tables(self.h, mg.ctypes.data_as(fp), cd.ctypes.data_as(fp), ms.ctypes.data_as(fp),
                                un.ctypes.data_as(bp))
        return mg, cd, ms, bool(un[0])

    def emissive_table(self):
        L = lib()
        m = L.oxrcpu_num_lights(self.h)
        tris = np.zeros(max(m, 1), np.uint32)
        pmf = np.zeros(max(m, 1), np.float32)
        cdf = np.zeros(max(m, 1), np.float32)
        lookup = np.zeros(max(self.scene.num_tris, 1), np.uint32)
        L.oxrcpu_emissive_table(self.h, tris.ctypes.data_as(up), pmf.ctypes.data_as(fp), cdf.ctypes.data_as(fp),
                                lookup.ctypes.data_as(up))
        return tris[:m], pmf[:m], cdf[:m], lookup[:self.scene.num_tris]

    def record_rays(self, width, height, first, n, y0, y1, x_step=1, max_rays=1 << 20, opts=None):
        """Closest-hit rays the oracle's paths trace (8 floats: o, k, d, t_max)."""
        if opts is None:
            opts = options(width, height, ambient=getattr(self.scene, "ambient", 0.0))
        cam = self.scene.camera.resolve(width, height)
        out = np.zeros((max_rays, 8), np.float32)
        m = lib().oxrcpu_record_rays(self.h, C.byref(cam), C.byref(opts), first, n, y0, y1, x_step, max_rays,
                                     out.ctypes.data_as(fp))
        return out[:m]

    def record_shadow_rays(self, width, height, first, n, y0, y1, x_step=1, max_rays=1 << 20, opts=None):
        """Occlusion rays the oracle's paths trace (8 floats: o, skip id bits,
        d, t_max)."""
        if opts is None:
            opts = options(width, height, ambient=getattr(self.scene, "ambient", 0.0))
        cam = self.scene.camera.resolve(width, height)
        out = np.zeros((max_rays, 8), np.float32)
        m = lib().oxrcpu_record_shadow_rays(self.h, C.byref(cam), C.byref(opts), first, n, y0, y1, x_step, max_rays,
                                            out.ctypes.data_as(fp))
        return out[:m]

    def tex_sample(self, tex, uv, lod=None, srgb=False):
        """PreparedTexture::sample_lod at (n, 2) uv and n levels (0 when
        lod is None): (n, 4) linear RGBA."""
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        lod = np.zeros(len(uv), np.float32) if lod is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(lod, np.float32), (len(uv),)))
        out = np.zeros((len(uv), 4), np.float32)
        if lib().oxrcpu_tex_sample(self.h, tex, 1 if srgb else 0, len(uv), uv.ctypes.data_as(fp),
                                   lod.ctypes.data_as(fp), out.ctypes.data_as(fp)) != 0:
            raise ValueError("no such texture")
        return out

    def tex_grad_lod(self, tex, duv):
        """sample_grad's level for (n, 4) (du/dx, dv/dx, du/dy, dv/dy)."""
        duv = np.ascontiguousarray(duv, np.float32).reshape(-1, 4)
        out = np.zeros(len(duv), np.float32)
        if lib().oxrcpu_tex_grad_lod(self.h, tex, len(duv), duv.ctypes.data_as(fp), out.ctypes.data_as(fp)) != 0:
            raise ValueError("no such texture")
        return out

    def tex_levels(self, tex, srgb=False):
        """The mip chain: a list of (h, w, 4) arrays."""
        wh = np.zeros(2, np.uint32)
        n = lib().oxrcpu_tex_level(self.h, tex, 1 if srgb else 0, 0, wh.ctypes.data_as(up), fp())
        if n < 0:
            raise ValueError("no such texture")
        out = []
        for lv in range(n):
            lib().oxrcpu_tex_level(self.h, tex, 1 if srgb else 0, lv, wh.ctypes.data_as(up), fp())
            t = np.zeros((int(wh[1]), int(wh[0]), 4), np.float32)
            lib().oxrcpu_tex_level(self.h, tex, 1 if srgb else 0, lv, wh.ctypes.data_as(up), t.ctypes.data_as(fp))
            out.append(t)
        return out

    def material_eval(self, g, bary, width=None, dirs=None, flip=None):
        """material_oriented at hits (triangle g, barycentrics): (n, 43)
        floats, see oxrcpu_material_eval."""
        g = np.ascontiguousarray(g, np.uint32).ravel()
        n = len(g)
        bary = np.ascontiguousarray(bary, np.float32).reshape(n, 3)
        width = np.zeros(n, np.float32) if width is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(width, np.float32), (n,)))
        dirs = np.zeros((n, 3), np.float32) if dirs is None else np.ascontiguousarray(dirs, np.float32).reshape(n, 3)
        flip = np.zeros(n, np.uint8) if flip is None else np.ascontiguousarray(flip, np.uint8).ravel()
        out = np.zeros((n, 43), np.float32)
        if lib().oxrcpu_material_eval(self.h, n, g.ctypes.data_as(up), bary.ctypes.data_as(fp),
                                      width.ctypes.data_as(fp), dirs.ctypes.data_as(fp), flip.ctypes.data_as(bp),
                                      out.ctypes.data_as(fp)) != 0:
            raise ValueError("bad triangle id")
        return out

    def surface(self, g, bary):
        b, bp_ = _f(bary)
        out = np.zeros(9, np.float32)
        lib().oxrcpu_surface(self.h, g, bp_, out.ctypes.data_as(fp))
        return out[0:3], out[3:6], out[6:9]


def rays_array(o, d, t_min=None, t_max=None):
    """Pack (n,3) origins/directions into an oxr_ray array."""
    o = np.asarray(o, np.float32).reshape(-1, 3)
    d = np.asarray(d, np.float32).reshape(-1, 3)
    n = len(o)
    buf = np.zeros((n, 8), np.float32)
    buf[:, 0:3] = o
    buf[:, 3] = 0.0 if t_min is None else t_min
    buf[:, 4:7] = d
    buf[:, 7] = np.inf if t_max is None else t_max
    arr = (A.Ray * max(n, 1))()
    C.memmove(arr, buf.ctypes.data, buf.nbytes)
    return arr, n


class RayBatch:
    def __init__(self, o, d, t_min=None, t_max=None):
        self.arr, self.n = rays_array(o, d, t_min, t_max)

    def __len__(self):
        return self.n


def hits_to_numpy(hits, n):
    raw = np.frombuffer(hits, dtype=np.uint32, count=4 * max(n, 1)).reshape(-1, 4)[:n].copy()
    return raw  # columns: t bits, prim, u bits, v bits


def sobol_owen_4d(idx, seed):
    idx = np.ascontiguousarray(idx, np.uint32)
    seed = np.ascontiguousarray(seed, np.uint32)
    out = np.zeros((len(idx), 4), np.float32)
    lib().oxrcpu_sobol_owen_4d(len(idx), idx.ctypes.data_as(up), seed.ctypes.data_as(up), out.ctypes.data_as(fp))
    return out


def sobol(index, dim):
    return lib().oxrcpu_sobol(index, dim)


def pcg_hash(x):
    return lib().oxrcpu_pcg_hash(x)


def sobol_matrices():
    out = np.zeros(128, np.uint32)
    lib().oxrcpu_sobol_matrices(out.ctypes.data_as(up))
    return out.reshape(4, 32)


def sheen_lut():
    out = np.zeros(512, np.float32)
    lib().oxrcpu_sheen_lut(out.ctypes.data_as(fp))
    return out


def _ms_array(ms):
    if isinstance(ms, A.MaterialSample):
        ms = [ms]
    return (A.MaterialSample * len(ms))(*ms)


def bsdf_eval(ms, v, ng, front, l):
    """Vectorised over the rows of l (v, ng, front, ms broadcast)."""
    l = np.ascontiguousarray(np.asarray(l, np.float32).reshape(-1, 3))
    n = len(l)
    if isinstance(ms, A.MaterialSample):
        ms = [ms] * n
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n, 3)))
    ng = np.ascontiguousarray(np.broadcast_to(np.asarray(ng, np.float32), (n, 3)))
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(front, np.uint8), (n,)))
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().oxrcpu_bsdf_eval(n, _ms_array(ms), v.ctypes.data_as(fp), ng.ctypes.data_as(fp), fr.ctypes.data_as(bp),
                           l.ctypes.data_as(fp), f.ctypes.data_as(fp), pdf.ctypes.data_as(fp))
    return f, pdf


def bsdf_sample(ms, v, ng, front, u):
    u = np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 3))
    n = len(u)
    if isinstance(ms, A.MaterialSample):
        ms = [ms] * n
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n, 3)))
    ng = np.ascontiguousarray(np.broadcast_to(np.asarray(ng, np.float32), (n, 3)))
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(front, np.uint8), (n,)))
    l = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    ok = np.zeros(n, np.uint8)
    lib().oxrcpu_bsdf_sample(n, _ms_array(ms), v.ctypes.data_as(fp), ng.ctypes.data_as(fp), fr.ctypes.data_as(bp),
                             u.ctypes.data_as(fp), l.ctypes.data_as(fp), f.ctypes.data_as(fp),
                             pdf.ctypes.data_as(fp), ok.ctypes.data_as(bp))
    return l, f, pdf, ok.astype(bool)


def bsdf_probs(ms, v, ng, front):
    v, vp = _f(v)
    ng, ngp = _f(ng)
    out = np.zeros(6, np.float32)
    lib().oxrcpu_bsdf_probs(C.byref(ms), vp, ngp, 1 if front else 0, out.ctypes.data_as(fp))
    return out


def triangle_solid_angle(p, a, b, c):
    args = [_f(x) for x in (p, a, b, c)]
    return lib().oxrcpu_triangle_solid_angle(*[x[1] for x in args])


def sample_spherical_triangle(p, a, b, c, u0, u1):
    args = [_f(x) for x in (p, a, b, c)]
    out = np.zeros(3, np.float32)
    ok = lib().oxrcpu_sample_spherical_triangle(*[x[1] for x in args], u0, u1, out.ctypes.data_as(fp))
    return out if ok else None


def iridescent_fresnel(outside, film, f0, thick, cos1):
    f0a, f0p = _f(f0)
    out = np.zeros(3, np.float32)
    lib().oxrcpu_iridescent_fresnel(outside, film, f0p, thick, cos1, out.ctypes.data_as(fp))
    return out


def offset_ray_origin(p, n):
    pa, pp = _f(p)
    na, np_ = _f(n)
    out = np.zeros(3, np.float32)
    lib().oxrcpu_offset_ray_origin(pp, np_, out.ctypes.data_as(fp))
    return out


def barycentric_of(q, p0, p1, p2):
    args = [_f(x) for x in (q, p0, p1, p2)]
    out = np.zeros(3, np.float32)
    ok = lib().oxrcpu_barycentric_of(*[x[1] for x in args], out.ctypes.data_as(fp))
    return out if ok else None


def resolve_hdr(accum, samples, background=(0, 0, 0, 255), tone_map=0, exposure=1.0):
    accum = np.ascontiguousarray(accum, np.float64)
    npx = accum.size // 4
    d = A.Display()
    d.background = (C.c_uint8 * 4)(*background)
    d.tone_map = tone_map
    d.exposure = exposure
    out = np.zeros(accum.shape[:-1] + (4,), np.float32)
    lib().oxrcpu_resolve_hdr(accum.ctypes.data_as(dp), npx, samples, C.byref(d), out.ctypes.data_as(fp))
    return out


def resolve_rgba8(accum, samples, background=(0, 0, 0, 255), tone_map=0, exposure=1.0):
    accum = np.ascontiguousarray(accum, np.float64)
    npx = accum.size // 4
    d = A.Display()
    d.background = (C.c_uint8 * 4)(*background)
    d.tone_map = tone_map
    d.exposure = exposure
    out = np.zeros(accum.shape[:-1] + (4,), np.uint8)
    lib().oxrcpu_resolve_rgba8(accum.ctypes.data_as(dp), npx, samples, C.byref(d), out.ctypes.data_as(bp))
    return out


def tone_map(rgb, tm=0):
    """ToneMap::apply (src/hdr.rs:44-66) on (n, 3) scene-linear colours."""
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros_like(rgb)
    lib().oxrcpu_tone_map(tm, len(rgb), rgb.ctypes.data_as(fp), out.ctypes.data_as(fp))
    return out


def linear_to_srgb_byte(c):
    """linear_to_srgb_byte (src/hdr.rs:116-129): threshold binary search."""
    c = np.ascontiguousarray(c, np.float32).ravel()
    out = np.zeros(len(c), np.uint8)
    lib().oxrcpu_linear_to_srgb_byte(len(c), c.ctypes.data_as(fp), out.ctypes.data_as(bp))
    return out


def srgb_to_linear(c):
    """srgb_to_linear (src/hdr.rs:76-82)."""
    c = np.ascontiguousarray(c, np.float32).ravel()
    out = np.zeros(len(c), np.float32)
    lib().oxrcpu_srgb_to_linear(len(c), c.ctypes.data_as(fp), out.ctypes.data_as(fp))
    return out


def bvh4_count(nodes, tris, rays):
    """(nodes fetched, triangles fetched) per ray on an exported device BVH."""
    rays = np.ascontiguousarray(rays, np.float32)
    counts = np.zeros((len(rays), 2), np.uint32)
    lib().oxrcpu_bvh4_count(nodes.ctypes.data, tris.ctypes.data, len(rays), rays.ctypes.data_as(fp),
                            counts.ctypes.data_as(up))
    return counts


def bvh4_count_any(nodes, tris, rays):
    """(nodes fetched, triangles tested) per occlusion ray (o, skip id bits,
    d, t_max) on an exported device BVH, in the device's any-hit order."""
    rays = np.ascontiguousarray(rays, np.float32)
    counts = np.zeros((len(rays), 2), np.uint32)
    lib().oxrcpu_bvh4_count_any(nodes.ctypes.data, tris.ctypes.data, len(rays), rays.ctypes.data_as(fp),
                                counts.ctypes.data_as(up))
    return counts


def bvh4_stack_depth(nodes, tris, rays):
    """Most traversal-stack entries each ray holds at once on an exported
    device BVH (closest-hit order)."""
    rays = np.ascontiguousarray(rays, np.float32)
    depth = np.zeros(len(rays), np.uint32)
    lib().oxrcpu_bvh4_stack_depth(nodes.ctypes.data, tris.ctypes.data, len(rays), rays.ctypes.data_as(fp),
                                  depth.ctypes.data_as(up))
    return depth


# ---------------------------------------------------------------------------
# Learned path guide (guide_cpu.cpp).
# ---------------------------------------------------------------------------
GUIDE_NIN = A.GUIDE_LEVELS * A.GUIDE_FEATURES


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def guide_num_grid(cfg):
    return A.GUIDE_LEVELS * (1 << cfg.log2_table) * A.GUIDE_FEATURES


def guide_resolutions(cfg):
    res = np.zeros(A.GUIDE_LEVELS, np.uint32)
    dense = np.zeros(A.GUIDE_LEVELS, np.uint8)
    lib().oxrcpu_guide_resolutions(C.byref(cfg), res.ctypes.data_as(up), dense.ctypes.data_as(bp))
    return res, dense.astype(bool)


def bf16_round(x):
    x = _f32(x)
    return np.array([lib().oxrcpu_bf16_round(float(v)) for v in x.ravel()], np.float32).reshape(x.shape)


def guide_corners(cfg, pos):
    pos = _f32(pos).reshape(-1, 3)
    n = len(pos)
    idx = np.zeros((n, A.GUIDE_LEVELS, 8), np.uint32)
    w = np.zeros((n, A.GUIDE_LEVELS, 8), np.float32)
    lib().oxrcpu_guide_corners(C.byref(cfg), n, pos.ctypes.data_as(fp), idx.ctypes.data_as(up), w.ctypes.data_as(fp))
    return idx, w


def guide_forward(cfg, grid, mlp, pos, mode=1):
    grid, mlp, pos = _f32(grid), _f32(mlp), _f32(pos).reshape(-1, 3)
    n = len(pos)
    feat = np.zeros((n, GUIDE_NIN), np.float32)
    raw = np.zeros((n, A.GUIDE_OUTPUTS), np.float32)
    lib().oxrcpu_guide_forward(C.byref(cfg), grid.ctypes.data_as(fp), mlp.ctypes.data_as(fp), n,
                               pos.ctypes.data_as(fp), mode, feat.ctypes.data_as(fp), raw.ctypes.data_as(fp))
    return feat, raw


def guide_lobes(raw):
    raw = _f32(raw).reshape(-1, A.GUIDE_OUTPUTS)
    out = np.zeros((len(raw), A.GUIDE_LOBES, 5), np.float32)
    lib().oxrcpu_guide_lobes(len(raw), raw.ctypes.data_as(fp), out.ctypes.data_as(fp))
    return out


def guide_pdf(lobes, dirs):
    lobes, dirs = _f32(lobes), _f32(dirs).reshape(-1, 3)
    out = np.zeros(len(dirs), np.float64)
    lib().oxrcpu_guide_pdf(len(dirs), lobes.ctypes.data_as(fp), dirs.ctypes.data_as(fp), out.ctypes.data_as(dp))
    return out


def guide_sample(lobes, u):
    lobes, u = _f32(lobes), _f32(u).reshape(-1, 3)
    out = np.zeros((len(u), 3), np.float32)
    lib().oxrcpu_guide_sample(len(u), lobes.ctypes.data_as(fp), u.ctypes.data_as(fp), out.ctypes.data_as(fp))
    return out


def guide_loss_grad(cfg, grid, mlp, pos, dirs, weight, mode=1, grads=True):
    grid, mlp = _f32(grid), _f32(mlp)
    pos, dirs, weight = _f32(pos).reshape(-1, 3), _f32(dirs).reshape(-1, 3), _f32(weight).ravel()
    gg = np.zeros(grid.size, np.float64) if grads else None
    gm = np.zeros(mlp.size, np.float64) if grads else None
    loss = lib().oxrcpu_guide_loss_grad(C.byref(cfg), grid.ctypes.data_as(fp), mlp.ctypes.data_as(fp), len(pos),
                                        pos.ctypes.data_as(fp), dirs.ctypes.data_as(fp), weight.ctypes.data_as(fp),
                                        mode, gg.ctypes.data_as(dp) if grads else dp(),
                                        gm.ctypes.data_as(dp) if grads else dp())
    return loss, gg, gm


def guide_adam(p, m, v, g, t, lr, b1, b2, eps, lazy):
    """In place on float32 arrays p, m, v."""
    g = _f32(g)
    lib().oxrcpu_guide_adam(p.size, p.ctypes.data_as(fp), m.ctypes.data_as(fp), v.ctypes.data_as(fp),
                            g.ctypes.data_as(fp), t, lr, b1, b2, eps, int(lazy))
