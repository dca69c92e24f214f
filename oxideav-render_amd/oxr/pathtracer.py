"""Python mirror of oxideav-render's PathTrace API over liboxr.so.

Names, argument meaning and error behaviour follow the reference:
`PathTraceOptions` / `RenderOptions` (src/options.rs:67-146, :326-414),
`PathTracer::{new, sync, refine, samples, is_converged, reset,
invalidate_scene, image, hdr}` (src/pathtrace.rs:2420-2735) and
`PathTraceRenderer::{render, render_hdr}` (:2745-2782). The scene handed
to `sync` is either a scene graph (`oxr.prepare.Scene3D`), which `sync`
prepares as the reference does (re-preparing when the animation time,
animation, variant or lights change), or an already prepared soup
(`oxr.scenes.Scene`).
"""
import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np

from . import _abi as A
from ._lib import BackendError, InvalidOptions, check, lib

MIS, LIGHT_ONLY, BSDF_ONLY = A.OXR_STRATEGY_MIS, A.OXR_STRATEGY_LIGHT_ONLY, A.OXR_STRATEGY_BSDF_ONLY
CLAMP, REINHARD, ACES = A.OXR_TONEMAP_CLAMP, A.OXR_TONEMAP_REINHARD, A.OXR_TONEMAP_ACES


@dataclass
class PathTraceOptions:
    samples_per_pixel: int = 64
    max_bounces: int = 8
    rr_start: int = 3
    clamp: float = 0.0
    seed: int = 0
    strategy: int = MIS


@dataclass
class RenderOptions:
    """The path tracer's part of RenderOptions (src/options.rs:326-414).
    The fields after `path_trace` apply to a scene graph handed to
    `PathTracer.sync`: its preparation (PrepareOptions::from_render_options,
    src/prepare.rs:137-146) and its camera (Camera::resolve: `scene_camera`,
    else auto-frame or orbit with `projection`, `fov_deg`, `camera` =
    (elevation_deg, azimuth_deg, distance) and `camera_target_offset`)."""
    width: int = 512
    height: int = 512
    background: tuple = (0, 0, 0, 0)
    ambient: float = 0.2
    tone_map: int = CLAMP
    exposure: float = 1.0
    path_trace: PathTraceOptions = field(default_factory=PathTraceOptions)
    time: float = None
    animation: int = None
    material_variant: int = None
    use_scene_lights: bool = True
    light: object = None             # prepare.LightSpec; None = its default
    scene_camera: int = None
    projection: int = A.OXR_PROJ_PERSPECTIVE
    fov_deg: float = 60.0
    camera: tuple = None
    camera_target_offset: tuple = (0.0, 0.0, 0.0)

    def validate(self):
        """RenderOptions::validate's path-trace checks (src/options.rs:723-741)."""
        pt = self.path_trace
        if not (1 <= pt.samples_per_pixel <= 1 << 20):
            raise InvalidOptions(f"path_trace.samples_per_pixel must be in 1..=1048576, got {pt.samples_per_pixel}")
        if pt.max_bounces > 1024:
            raise InvalidOptions(f"path_trace.max_bounces must be <= 1024, got {pt.max_bounces}")
        if not (np.isfinite(pt.clamp) and pt.clamp >= 0.0):
            raise InvalidOptions(f"path_trace.clamp must be finite and >= 0.0, got {pt.clamp}")
        if self.width < 1 or self.height < 1:
            raise InvalidOptions("width and height must be >= 1")
        if not (np.isfinite(self.ambient) and self.ambient >= 0.0):
            raise InvalidOptions(f"ambient must be finite and >= 0, got {self.ambient}")
        return self


def _pt_opts(o: RenderOptions):
    p = A.PtOptions()
    p.width, p.height = max(o.width, 1), max(o.height, 1)
    pt = o.path_trace
    p.samples_per_pixel = pt.samples_per_pixel
    p.max_bounces = pt.max_bounces
    p.rr_start = pt.rr_start
    p.clamp = pt.clamp
    p.seed = pt.seed & 0xFFFFFFFF
    p.strategy = pt.strategy
    p.ambient = o.ambient
    return p


def _display(o: RenderOptions):
    d = A.Display()
    d.background = (C.c_uint8 * 4)(*o.background)
    d.tone_map = o.tone_map
    d.exposure = o.exposure
    return d


class PathTracer:
    """Progressive path tracer on one MI355X: scene, BVH, light tables and
    the f64 accumulator stay resident in HBM between refines."""

    def __init__(self, device=0):
        self._L = lib()
        h = C.c_void_p()
        check(self._L.oxr_pt_create(device, C.byref(h)), "oxr_pt_create")
        self._h = h
        self._device = device
        self._scene = None
        self._graph = None      # the scene graph the uploaded soup was prepared from
        self._prep_key = None
        self._prepared = None
        self._desc = None
        self._dirty = True
        self._opts = RenderOptions()

    def close(self):
        if getattr(self, "_h", None):
            self._L.oxr_pt_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # -- PathTracer API ------------------------------------------------------
    def invalidate_scene(self):
        self._dirty = True

    def reset(self):
        check(self._L.oxr_pt_reset(self._h), "reset")

    def set_window(self, x0=0, y0=0, w=0, h=0):
        """Trace only frame pixels [x0, x0 + w) x [y0, y0 + h), with the
        samples they have in the whole-frame render; accum / hdr / image
        then hold just that window. w = h = 0: the whole frame. Resets."""
        check(self._L.oxr_pt_set_window(self._h, x0, y0, w, h), "set_window")
        self._win = (x0, y0, w, h) if w else None

    def _dims(self):
        if getattr(self, "_win", None):
            return self._win[2], self._win[3]
        return max(self._opts.width, 1), max(self._opts.height, 1)

    def sync(self, scene, opts: RenderOptions):
        """Bring the tracer in line with `scene` + `opts`; returns True when
        the accumulation was reset (src/pathtrace.rs:2474-2508). A scene
        graph is prepared again when this tracer is dirty, the graph is
        another object, or its preparation inputs (time, animation,
        variant, scene lights, fallback light) changed."""
        from . import prepare as P
        reset = False
        if isinstance(scene, P.Scene3D):
            prep_key = (opts.time if opts.time is not None and np.isfinite(opts.time) else None, opts.animation,
                        opts.material_variant, bool(opts.use_scene_lights),
                        opts.light if opts.light is not None else P.LightSpec())
            if self._dirty or self._graph is not scene or self._prep_key != prep_key:
                self._prepared = P.build(scene, use_scene_lights=prep_key[3], fallback_light=prep_key[4],
                                         time=prep_key[0], animation=opts.animation,
                                         material_variant=opts.material_variant)
                soup = P.to_scene(self._prepared)
                self._desc = soup.desc()
                check(self._L.oxr_pt_upload_scene(self._h, C.byref(self._desc)), "upload_scene")
                self._scene, self._graph, self._prep_key = soup, scene, prep_key
                self._dirty = False
                reset = True
            cam = P.PreparedCamera(self._prepared, opts.scene_camera, opts.projection, opts.fov_deg, opts.camera,
                                   opts.camera_target_offset).resolve(max(opts.width, 1), max(opts.height, 1))
        else:
            if self._dirty or self._scene is not scene:
                self._desc = scene.desc()
                check(self._L.oxr_pt_upload_scene(self._h, C.byref(self._desc)), "upload_scene")
                self._scene, self._graph = scene, None
                self._dirty = False
                reset = True
            cam = scene.camera.resolve(max(opts.width, 1), max(opts.height, 1))
        r1 = C.c_int(0)
        r2 = C.c_int(0)
        check(self._L.oxr_pt_set_options(self._h, C.byref(_pt_opts(opts)), C.byref(r1)), "set_options")
        check(self._L.oxr_pt_set_camera(self._h, C.byref(cam), C.byref(r2)), "set_camera")
        self._opts = replace(opts)
        return bool(reset or r1.value or r2.value)

    def refine(self, samples):
        check(self._L.oxr_pt_refine(self._h, int(samples) & 0xFFFFFFFF), "refine")

    def samples(self):
        n = C.c_uint32(0)
        check(self._L.oxr_pt_samples(self._h, C.byref(n)), "samples")
        return n.value

    def is_converged(self):
        return self.samples() >= self._opts.path_trace.samples_per_pixel

    def accum(self, with_sum_sq=False):
        """The accumulator: (H, W, 4) f64 sums (r, g, b, covered count)
        [and (H, W, 3) squared radiance]."""
        W, H = self._dims()
        acc = np.zeros((H, W, 4), np.float64)
        sq = np.zeros((H, W, 3), np.float64) if with_sum_sq else None
        dp = C.POINTER(C.c_double)
        check(self._L.oxr_pt_read_accum(self._h, acc.ctypes.data_as(dp),
                                        sq.ctypes.data_as(dp) if sq is not None else dp()), "read_accum")
        return (acc, sq) if with_sum_sq else acc

    def hdr(self):
        W, H = self._dims()
        out = np.zeros((H, W, 4), np.float32)
        check(self._L.oxr_pt_resolve_hdr(self._h, C.byref(_display(self._opts)),
                                         out.ctypes.data_as(C.POINTER(C.c_float))), "resolve_hdr")
        return out

    def image(self):
        W, H = self._dims()
        out = np.zeros((H, W, 4), np.uint8)
        check(self._L.oxr_pt_resolve_rgba8(self._h, C.byref(_display(self._opts)),
                                           out.ctypes.data_as(C.POINTER(C.c_uint8))), "resolve_rgba8")
        return out

    # -- environment map ------------------------------------------------------
    def set_environment(self, rgba, intensity=1.0):
        """PathTracer::set_environment (src/pathtrace.rs:2449-2454): an
        equirectangular HDR image (H, W, 4), scene-linear, row 0 = +Y, scaled
        by `intensity`; None removes it. Replaces the ambient sky; resets."""
        fpp = C.POINTER(C.c_float)
        if rgba is None:
            self._env = None
            check(self._L.oxr_pt_set_environment(self._h, fpp(), 0, 0, 0.0), "set_environment")
            return
        img = np.ascontiguousarray(rgba, dtype=np.float32)
        if img.ndim != 3 or img.shape[2] != 4:
            raise InvalidOptions("environment image must be (H, W, 4)")
        self._env = img
        check(self._L.oxr_pt_set_environment(self._h, img.ctypes.data_as(fpp), img.shape[1], img.shape[0],
                                             float(intensity)), "set_environment")

    def tex_sample(self, tex, uv, lod=None, srgb=False):
        """sample_lod of texture `tex` (in the colour space a material
        uses it in) at (n, 2) uv and n levels: (n, 4)."""
        fpp = C.POINTER(C.c_float)
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        lod = np.zeros(len(uv), np.float32) if lod is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(lod, np.float32), (len(uv),)))
        out = np.zeros((len(uv), 4), np.float32)
        check(self._L.oxr_pt_tex_sample(self._h, tex, 1 if srgb else 0, len(uv), uv.ctypes.data_as(fpp),
                                        lod.ctypes.data_as(fpp), out.ctypes.data_as(fpp)), "tex_sample")
        return out

    def material_eval(self, g, bary, width=None, dirs=None, flip=None):
        """material_oriented + emission + alpha at hits: (n, 43)."""
        fpp = C.POINTER(C.c_float)
        g = np.ascontiguousarray(g, np.uint32).ravel()
        n = len(g)
        bary = np.ascontiguousarray(bary, np.float32).reshape(n, 3)
        width = np.zeros(n, np.float32) if width is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(width, np.float32), (n,)))
        dirs = np.zeros((n, 3), np.float32) if dirs is None else np.ascontiguousarray(dirs, np.float32).reshape(n, 3)
        flip = np.zeros(n, np.uint8) if flip is None else np.ascontiguousarray(flip, np.uint8).ravel()
        out = np.zeros((n, 43), np.float32)
        check(self._L.oxr_pt_material_eval(self._h, n, g.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           bary.ctypes.data_as(fpp), width.ctypes.data_as(fpp),
                                           dirs.ctypes.data_as(fpp), flip.ctypes.data_as(C.POINTER(C.c_uint8)),
                                           out.ctypes.data_as(fpp)), "material_eval")
        return out

    def env_eval(self, dirs):
        fpp = C.POINTER(C.c_float)
        d = np.ascontiguousarray(dirs, dtype=np.float32).reshape(-1, 3)
        pdf = np.zeros(len(d), np.float32)
        le = np.zeros((len(d), 3), np.float32)
        check(self._L.oxr_pt_env_eval(self._h, len(d), d.ctypes.data_as(fpp), pdf.ctypes.data_as(fpp),
                                      le.ctypes.data_as(fpp)), "env_eval")
        return pdf, le

    def env_sample(self, u):
        fpp = C.POINTER(C.c_float)
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1, 2)
        d = np.zeros((len(u), 3), np.float32)
        pdf = np.zeros(len(u), np.float32)
        ok = np.zeros(len(u), np.uint8)
        check(self._L.oxr_pt_env_sample(self._h, len(u), u.ctypes.data_as(fpp), d.ctypes.data_as(fpp),
                                        pdf.ctypes.data_as(fpp), ok.ctypes.data_as(C.POINTER(C.c_uint8))),
              "env_sample")
        return d, pdf, ok.astype(bool)

    # -- learned path guide ----------------------------------------------------
    def set_guide(self, guide, train=True):
        """Attach an oxr.guide.Guide (None detaches); resets the
        accumulation. With train=True every refine pass also trains it."""
        self._guide = guide  # keep it alive while attached
        check(self._L.oxr_pt_set_guide(self._h, guide.handle if guide is not None else None, 1 if train else 0),
              "set_guide")

    # -- instrumentation -----------------------------------------------------
    def set_timing(self, enable=True, count=False):
        """enable: time every kernel with HIP events; count: count the BVH
        nodes and triangles closest-hit rays fetch (stats closest_nodes /
        closest_tris). Either runs the bounce loop without graphs."""
        check(self._L.oxr_pt_set_timing(self._h, (1 if enable else 0) | (2 if count else 0)), "set_timing")

    def stats(self):
        s = A.PtStats()
        check(self._L.oxr_pt_stats_get(self._h, C.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in A.PtStats._fields_}

    def autotune(self, spp=4):
        """Time refine(spp) under each candidate launch configuration, keep
        the fastest; resets the accumulation. Returns the choice."""
        r = A.TuneResult()
        check(self._L.oxr_pt_autotune(self._h, int(spp), C.byref(r)), "autotune")
        return {f: getattr(r, f) for f, _ in A.TuneResult._fields_}

    def bvh_info(self):
        b = A.BvhInfo()
        check(self._L.oxr_pt_bvh_info(self._h, C.byref(b)), "bvh_info")
        return {f: getattr(b, f) for f, _ in A.BvhInfo._fields_}

    def bvh_export(self):
        info = self.bvh_info()
        nodes = np.zeros(info["num_nodes"] * 16, np.uint32)
        tris = np.zeros(max(info["num_tris"], 1) * 12, np.uint32)
        check(self._L.oxr_pt_bvh_export(self._h, nodes.ctypes.data, tris.ctypes.data), "bvh_export")
        return nodes.reshape(-1, 16), tris.reshape(-1, 12)[:info["num_tris"]]

    # -- parity entry points -------------------------------------------------
    def upload(self, scene):
        """Upload a scene without options (for the trace entry points)."""
        self._desc = scene.desc()
        check(self._L.oxr_pt_upload_scene(self._h, C.byref(self._desc)), "upload_scene")
        self._scene, self._graph = scene, None
        self._dirty = False

    def trace_closest(self, rays, n, flags=0):
        hits = (A.Hit * max(n, 1))()
        check(self._L.oxr_trace_closest(self._h, n, rays, flags, hits), "trace_closest")
        return np.frombuffer(hits, dtype=np.uint32, count=4 * max(n, 1)).reshape(-1, 4)[:n].copy()

    def trace_any(self, rays, n):
        occ = np.zeros(max(n, 1), np.uint8)
        check(self._L.oxr_trace_any(self._h, n, rays, occ.ctypes.data_as(C.POINTER(C.c_uint8))), "trace_any")
        return occ[:n]


class PathTraceRenderer:
    """`Renderer` over PathTracer: every render re-prepares the scene and
    takes samples_per_pixel samples (src/pathtrace.rs:2745-2782)."""

    def __init__(self, device=0):
        self.tracer = PathTracer(device)

    def _run(self, scene, opts):
        self.tracer.invalidate_scene()
        self.tracer.sync(scene, opts)
        self.tracer.refine(max(opts.path_trace.samples_per_pixel, 1))

    def render(self, scene, opts):
        self._run(scene, opts)
        return self.tracer.image()

    def render_hdr(self, scene, opts):
        self._run(scene, opts)
        return self.tracer.hdr()


def make_renderer(backend="pathtrace", device=0):
    """make_renderer / the registry name (src/lib.rs:174-180,
    src/registry.rs:110-113); this library provides the PathTrace backend."""
    if backend in ("pathtrace", "pathtrace-hip"):
        return PathTraceRenderer(device)
    raise BackendError(f"renderer backend '{backend}' not registered")


def sobol_owen_4d(index, pattern_seed, device=0):
    index = np.ascontiguousarray(index, np.uint32)
    seed = np.ascontiguousarray(pattern_seed, np.uint32)
    out = np.zeros((len(index), 4), np.float32)
    up = C.POINTER(C.c_uint32)
    check(lib().oxr_sobol_owen_4d(device, len(index), index.ctypes.data_as(up), seed.ctypes.data_as(up),
                                  out.ctypes.data_as(C.POINTER(C.c_float))), "sobol_owen_4d")
    return out


def _bsdf_args(ms, n):
    if isinstance(ms, A.MaterialSample):
        ms = [ms] * n
    return (A.MaterialSample * n)(*ms)


def bsdf_eval(ms, v, ng, front, l, device=0):
    l = np.ascontiguousarray(np.asarray(l, np.float32).reshape(-1, 3))
    n = len(l)
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n, 3)))
    ng = np.ascontiguousarray(np.broadcast_to(np.asarray(ng, np.float32), (n, 3)))
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(front, np.uint8), (n,)))
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    fp = C.POINTER(C.c_float)
    check(lib().oxr_bsdf_eval(device, n, _bsdf_args(ms, n), v.ctypes.data_as(fp), ng.ctypes.data_as(fp),
                              fr.ctypes.data_as(C.POINTER(C.c_uint8)), l.ctypes.data_as(fp), f.ctypes.data_as(fp),
                              pdf.ctypes.data_as(fp)), "bsdf_eval")
    return f, pdf


def bsdf_sample(ms, v, ng, front, u, device=0):
    u = np.ascontiguousarray(np.asarray(u, np.float32).reshape(-1, 3))
    n = len(u)
    v = np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n, 3)))
    ng = np.ascontiguousarray(np.broadcast_to(np.asarray(ng, np.float32), (n, 3)))
    fr = np.ascontiguousarray(np.broadcast_to(np.asarray(front, np.uint8), (n,)))
    l = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    ok = np.zeros(n, np.uint8)
    fp = C.POINTER(C.c_float)
    bp = C.POINTER(C.c_uint8)
    check(lib().oxr_bsdf_sample(device, n, _bsdf_args(ms, n), v.ctypes.data_as(fp), ng.ctypes.data_as(fp),
                                fr.ctypes.data_as(bp), u.ctypes.data_as(fp), l.ctypes.data_as(fp),
                                f.ctypes.data_as(fp), pdf.ctypes.data_as(fp), ok.ctypes.data_as(bp)), "bsdf_sample")
    return l, f, pdf, ok.astype(bool)
