"""Procedural scenes in the prepared-soup data contract.

A `Scene` is what `PathTracer::sync` hands the trace layer after
`PreparedScene::build` (oxideav-render src/prepare.rs:1098-1234,
src/trace.rs:320-351): a world-space, de-indexed triangle soup (CCW
front), optional per-corner normals, one material index per triangle, the
material table and the punctual lights. The geometry builders restate the
reference's test vocabulary (`quad`, `cuboid`, `uv_sphere`,
src/testscenes.rs:38-148); the scenes are the benchmark workloads:

* `cornell()`        — Cornell box, 36 tris, emissive ceiling quad.
* `material_sweep()` — 7x7 UV spheres (metallic rows x roughness columns)
                        on a checker floor, point + spot light, sky 0.3.
* `glossy_caustic()` — room, glass sphere, small bright emitter, GGX floor.
* `triangle_soup()`  — seeded random soup (serving mix).
* small test scenes: furnace box, area-light floor, glass cube, point-lit
  plane (the reference's own integrator tests, src/pathtrace.rs:3346-3750).

All scene randomness is `pcg_hash` chained from a stated seed.
"""
import ctypes as C
import math

import numpy as np

from . import _abi as A

F32 = np.float32
MATERIAL_DTYPE = np.dtype(A.Material)


def material_array(n, template):
    """n copies of the A.Material `template` as a structured array."""
    return np.repeat(np.frombuffer(bytes(template), dtype=MATERIAL_DTYPE), n)


def _normalize(v):
    v = np.asarray(v, dtype=F32)
    n = F32(math.sqrt(float(np.dot(v, v))))
    return v / n if n > 0 else v


class CameraSpec:
    """A glTF-style perspective/orthographic camera placed by look-at.

    `resolve(width, height)` follows Camera::from_scene_camera
    (src/camera.rs:208-332): forward, side = forward x up_hint, up = side x
    forward, half_h = tan(yfov/2), half_w = half_h * aspect (viewport
    aspect), in f32.
    """

    def __init__(self, eye, target, up=(0.0, 1.0, 0.0), yfov_deg=40.0, projection=0, ymag=1.0):
        self.eye = np.asarray(eye, dtype=F32)
        self.target = np.asarray(target, dtype=F32)
        self.up_hint = np.asarray(up, dtype=F32)
        self.yfov = F32(math.radians(yfov_deg))
        self.projection = projection
        self.ymag = F32(ymag)

    def resolve(self, width, height):
        fwd = _normalize(self.target - self.eye)
        side = _normalize(np.cross(fwd, _normalize(self.up_hint)).astype(F32))
        up = np.cross(side, fwd).astype(F32)
        aspect = F32(max(width, 1)) / F32(max(height, 1))
        cam = A.Camera()
        cam.eye = A.f3(*self.eye.tolist())
        cam.forward = A.f3(*fwd.tolist())
        cam.side = A.f3(*side.tolist())
        cam.up = A.f3(*up.tolist())
        if self.projection == A.OXR_PROJ_ORTHOGRAPHIC:
            cam.half_h = float(self.ymag)
            cam.half_w = float(F32(self.ymag) * aspect)
        else:
            th = F32(math.tan(float(self.yfov * F32(0.5))))
            cam.half_h = float(th)
            cam.half_w = float(th * aspect)
        cam.projection = self.projection
        return cam


class Scene:
    """Prepared triangle soup + materials + lights + a camera."""

    def __init__(self, name):
        self.name = name
        self._pos = []
        self._nrm = []
        self._has_n = []
        self._mid = []
        self._attr = {"uv0": [], "uv1": [], "tangents": [], "colors": []}
        self.materials = []
        self.lights = []
        self.textures = []         # (A.Texture, pixel array) pairs
        self.material_tex = {}     # material index -> A.MaterialTextures
        self.camera = None
        self._packed = None

    # -- building ----------------------------------------------------------
    def add_material(self, m):
        self.materials.append(m)
        return self.num_materials - 1

    def add_material_array(self, arr):
        """Append a numpy array of dtype MATERIAL_DTYPE; returns the index
        of its first entry."""
        first = self.num_materials
        self.materials.append(np.ascontiguousarray(arr, dtype=MATERIAL_DTYPE))
        return first

    @property
    def num_materials(self):
        return sum(len(m) if isinstance(m, np.ndarray) else 1 for m in self.materials)

    def material_table(self):
        """All materials as one MATERIAL_DTYPE array."""
        parts = [m if isinstance(m, np.ndarray) else np.frombuffer(bytes(m), dtype=MATERIAL_DTYPE)
                 for m in self.materials]
        if not parts:
            return np.zeros(0, MATERIAL_DTYPE)
        return np.ascontiguousarray(np.concatenate(parts))

    def material(self, i):
        """Material i as an A.Material."""
        t = self.material_table()
        return A.Material.from_buffer_copy(t[i:i + 1].tobytes())

    _ATTR = {"uv0": (6, A.TRI_HAS_UV0), "uv1": (6, A.TRI_HAS_UV1), "tangents": (12, A.TRI_HAS_TANGENTS),
             "colors": (12, A.TRI_HAS_COLORS)}

    def add_tris(self, pos, mat, normals=None, uv0=None, uv1=None, tangents=None, colors=None):
        """pos: (n, 9) corners; normals: (n, 9) or None (geometric);
        uv0 / uv1: (n, 6) TEXCOORD_0/1 per corner; tangents: (n, 12) xyzw
        per corner; colors: (n, 12) COLOR_0 RGBA per corner."""
        pos = np.ascontiguousarray(pos, dtype=F32).reshape(-1, 9)
        self._pos.append(pos)
        flags = np.zeros(len(pos), dtype=np.uint32)
        if normals is None:
            self._nrm.append(np.zeros_like(pos))
        else:
            self._nrm.append(np.ascontiguousarray(normals, dtype=F32).reshape(-1, 9))
            flags |= A.OXR_TRI_HAS_NORMALS
        for name, arr in (("uv0", uv0), ("uv1", uv1), ("tangents", tangents), ("colors", colors)):
            width, bit = self._ATTR[name]
            if arr is None:
                self._attr[name].append(np.zeros((len(pos), width), F32))
            else:
                self._attr[name].append(np.ascontiguousarray(arr, dtype=F32).reshape(len(pos), width))
                flags |= bit
        self._has_n.append(flags)
        if np.isscalar(mat):
            self._mid.append(np.full(len(pos), mat, dtype=np.uint32))
        else:
            self._mid.append(np.asarray(mat, dtype=np.uint32))
        self._packed = None

    def add_texture(self, pixels, mag=A.FILTER_LINEAR, min_filter=A.MIN_LINEAR_MIP_LINEAR, wrap_s=A.WRAP_REPEAT,
                    wrap_t=A.WRAP_REPEAT):
        """pixels: (H, W, 4) uint8 (RGBA8) or float32 (linear RGBA); row 0
        at v = 0. Returns the texture index."""
        px = np.ascontiguousarray(pixels)
        if px.ndim != 3 or px.shape[2] != 4:
            raise ValueError("texture pixels must be (H, W, 4)")
        if px.dtype == np.uint8:
            fmt = A.TEXFMT_RGBA8
        else:
            px = np.ascontiguousarray(px, dtype=F32)
            fmt = A.TEXFMT_RGBA32F
        t = A.Texture()
        t.width, t.height, t.format = px.shape[1], px.shape[0], fmt
        t.pixels = px.ctypes.data
        t.mag_filter, t.min_filter, t.wrap_s, t.wrap_t = mag, min_filter, wrap_s, wrap_t
        self.textures.append((t, px))
        return len(self.textures) - 1

    def set_material_textures(self, mat, normal_scale=1.0, clearcoat_normal_scale=1.0,
                              iridescence_thickness_min=100.0, **slots):
        """Bind textures to material `mat`. Each slot (base_color,
        metallic_roughness, normal, emissive, or a KHR extension slot) takes
        a texture index or a dict(texture=, uv_set=0, offset=(0, 0),
        rotation=0, scale=(1, 1)); giving offset, rotation or scale turns
        the KHR_texture_transform on."""
        m = A.material_textures()
        m.normal_scale, m.clearcoat_normal_scale = normal_scale, clearcoat_normal_scale
        m.iridescence_thickness_min = iridescence_thickness_min
        for name, v in slots.items():
            if name not in A.TEXTURE_SLOTS:
                raise ValueError(f"unknown texture slot {name!r}")
            b = A.no_binding()
            if isinstance(v, dict):
                b.texture = int(v["texture"])
                b.uv_set = int(v.get("uv_set", 0))
                if any(k in v for k in ("offset", "rotation", "scale")):
                    b.has_transform = 1
                    b.offset[0], b.offset[1] = v.get("offset", (0.0, 0.0))
                    b.rotation = float(v.get("rotation", 0.0))
                    b.scale[0], b.scale[1] = v.get("scale", (1.0, 1.0))
            else:
                b.texture = int(v)
            setattr(m, name, b)
        self.material_tex[int(mat)] = m
        self._packed = None

    def pack(self):
        if self._packed is None:
            if self._pos:
                pos = np.ascontiguousarray(np.concatenate(self._pos), dtype=F32)
                nrm = np.ascontiguousarray(np.concatenate(self._nrm), dtype=F32)
                hn = np.ascontiguousarray(np.concatenate(self._has_n), dtype=np.uint32)
                mid = np.ascontiguousarray(np.concatenate(self._mid), dtype=np.uint32)
                attr = {k: np.ascontiguousarray(np.concatenate(v), dtype=F32) for k, v in self._attr.items()}
            else:
                pos = np.zeros((0, 9), F32)
                nrm = np.zeros((0, 9), F32)
                hn = np.zeros(0, np.uint32)
                mid = np.zeros(0, np.uint32)
                attr = {k: np.zeros((0, self._ATTR[k][0]), F32) for k in self._attr}
            self._packed = (pos, nrm, hn, mid)
            self._packed_attr = attr
        return self._packed

    def attributes(self):
        """Per-triangle uv0, uv1, tangents and colors arrays (zeros where a
        triangle has none; see the tri flags)."""
        self.pack()
        return self._packed_attr

    @property
    def num_tris(self):
        return len(self.pack()[0])

    @property
    def positions(self):
        return self.pack()[0]

    @property
    def material_ids(self):
        return self.pack()[3]

    def desc(self):
        """oxr_scene_desc pointing at this scene's arrays (kept alive by
        the returned object)."""
        pos, nrm, hn, mid = self.pack()
        attr = self.attributes()
        d = A.SceneDesc()
        d.num_tris = len(pos)
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_uint32)
        d.positions = pos.ctypes.data_as(fp)
        any_n = bool((hn & A.OXR_TRI_HAS_NORMALS).any())
        d.normals = nrm.ctypes.data_as(fp) if any_n else fp()
        d.tri_flags = hn.ctypes.data_as(up) if bool(hn.any()) else up()
        for name, (_, bit) in self._ATTR.items():
            setattr(d, name, attr[name].ctypes.data_as(fp) if bool((hn & bit).any()) else fp())
        d.material_ids = mid.ctypes.data_as(up)
        mats = self.material_table()
        if len(mats) == 0:
            mats = np.zeros(1, MATERIAL_DTYPE)
            d.num_materials = 0
        else:
            d.num_materials = len(mats)
        lights = (A.Light * max(len(self.lights), 1))(*self.lights)
        d.materials = mats.ctypes.data_as(C.POINTER(A.Material))
        d.num_lights = len(self.lights)
        d.lights = C.cast(lights, C.POINTER(A.Light))
        texs = (A.Texture * max(len(self.textures), 1))(*[t for t, _ in self.textures])
        d.num_textures = len(self.textures)
        d.textures = C.cast(texs, C.POINTER(A.Texture))
        mtex = None
        if self.material_tex:
            mtex = (A.MaterialTextures * max(d.num_materials, 1))()
            for i in range(d.num_materials):
                mtex[i] = self.material_tex.get(i, A.material_textures())
            d.material_textures = C.cast(mtex, C.POINTER(A.MaterialTextures))
        d._keep = (pos, nrm, hn, mid, mats, lights, attr, texs, mtex, [px for _, px in self.textures])
        return d

    def bounds(self):
        p = self.positions.reshape(-1, 3)
        return p.min(axis=0), p.max(axis=0)


# ---------------------------------------------------------------------------
# Geometry builders (src/testscenes.rs:38-148, restated as soups).
# ---------------------------------------------------------------------------

def quad(p):
    """Planar quad p0->p1->p2->p3 (CCW from the front): triangles
    (0,1,2), (0,2,3) with the flat normal of cross(p1-p0, p3-p0)."""
    p = np.asarray(p, dtype=F32)
    n = _normalize(np.cross(p[1] - p[0], p[3] - p[0]).astype(F32))
    tris = np.stack([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])])
    nrm = np.tile(np.concatenate([n, n, n]), (2, 1))
    return tris.astype(F32), nrm.astype(F32)


def cuboid(mn, mx):
    """Axis-aligned box with outward faces (12 triangles)."""
    x0, y0, z0 = mn
    x1, y1, z1 = mx
    faces = [
        [[x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]],
        [[x1, y0, z0], [x0, y0, z0], [x0, y1, z0], [x1, y1, z0]],
        [[x1, y0, z1], [x1, y0, z0], [x1, y1, z0], [x1, y1, z1]],
        [[x0, y0, z0], [x0, y0, z1], [x0, y1, z1], [x0, y1, z0]],
        [[x0, y1, z1], [x1, y1, z1], [x1, y1, z0], [x0, y1, z0]],
        [[x0, y0, z0], [x1, y0, z0], [x1, y0, z1], [x0, y0, z1]],
    ]
    t, n = zip(*(quad(f) for f in faces))
    return np.concatenate(t), np.concatenate(n)


def uv_sphere(radius, segments, rings):
    """UV sphere (src/testscenes.rs:106-142): smooth normals; the pole
    rows keep their zero-area triangles, as the reference's builder does."""
    segments = max(segments, 3)
    rings = max(rings, 2)
    r_ = np.arange(rings + 1, dtype=F32)
    s_ = np.arange(segments + 1, dtype=F32)
    v = r_ / F32(rings)
    u = s_ / F32(segments)
    theta = (v * F32(math.pi)).astype(F32)
    phi = (u * F32(2.0) * F32(math.pi)).astype(F32)
    st, ct = np.sin(theta).astype(F32), np.cos(theta).astype(F32)
    sp, cp = np.sin(phi).astype(F32), np.cos(phi).astype(F32)
    n = np.stack([st[:, None] * cp[None, :], np.broadcast_to(ct[:, None], (rings + 1, segments + 1)),
                  -st[:, None] * sp[None, :]], axis=-1).astype(F32).reshape(-1, 3)
    pos = (n * F32(radius)).astype(F32)
    row = segments + 1
    a = (np.arange(rings)[:, None] * row + np.arange(segments)[None, :]).reshape(-1)
    b = a + row
    idx = np.stack([a, b, a + 1, a + 1, b, b + 1], axis=1).reshape(-1, 3)
    return pos[idx].reshape(-1, 9), n[idx].reshape(-1, 9)


def uv_sphere_textured(radius, segments, rings, center=(0.0, 0.0, 0.0)):
    """uv_sphere's grid with per-corner TEXCOORD_0 (u along phi, v along
    theta) and tangents (d/dphi, w = 1): (pos, nrm, uv, tangents)."""
    r_ = np.arange(rings + 1, dtype=F32)
    s_ = np.arange(segments + 1, dtype=F32)
    v = r_ / F32(rings)
    u = s_ / F32(segments)
    theta = (v * F32(math.pi)).astype(F32)
    phi = (u * F32(2.0) * F32(math.pi)).astype(F32)
    st, ct = np.sin(theta).astype(F32), np.cos(theta).astype(F32)
    sp, cp = np.sin(phi).astype(F32), np.cos(phi).astype(F32)
    n = np.stack([st[:, None] * cp[None, :], np.broadcast_to(ct[:, None], (rings + 1, segments + 1)),
                  -st[:, None] * sp[None, :]], axis=-1).astype(F32).reshape(-1, 3)
    pos = (n * F32(radius) + np.asarray(center, F32)).astype(F32)
    uv = np.stack(np.broadcast_arrays(u[None, :], v[:, None]), axis=-1).astype(F32).reshape(-1, 2)
    tg = np.stack([np.broadcast_to(-sp[None, :], (rings + 1, segments + 1)), np.zeros((rings + 1, segments + 1), F32),
                   np.broadcast_to(-cp[None, :], (rings + 1, segments + 1)),
                   np.ones((rings + 1, segments + 1), F32)], axis=-1).astype(F32).reshape(-1, 4)
    row = segments + 1
    a = (np.arange(rings)[:, None] * row + np.arange(segments)[None, :]).reshape(-1)
    b = a + row
    idx = np.stack([a, b, a + 1, a + 1, b, b + 1], axis=1).reshape(-1, 3)
    return pos[idx].reshape(-1, 9), n[idx].reshape(-1, 9), uv[idx].reshape(-1, 6), tg[idx].reshape(-1, 12)


def quad_uv(p, uv=((0, 0), (1, 0), (1, 1), (0, 1))):
    """quad() with per-corner TEXCOORD_0 of its corners p0..p3."""
    t, nrm = quad(p)
    uv = np.asarray(uv, F32)
    uvs = np.stack([np.concatenate([uv[0], uv[1], uv[2]]), np.concatenate([uv[0], uv[2], uv[3]])]).astype(F32)
    return t, nrm, uvs


def rotate_y(points9, deg, center):
    """Rotate (n, 9) corners about the Y axis through `center`."""
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], dtype=np.float64)
    p = points9.reshape(-1, 3).astype(np.float64) - np.asarray(center, np.float64)
    return ((p @ R.T) + np.asarray(center, np.float64)).astype(F32).reshape(-1, 9)


def translate(points9, t):
    return (points9.reshape(-1, 3) + np.asarray(t, dtype=F32)).astype(F32).reshape(-1, 9)


def lambert(rgb, **kw):
    """Lambertian material: dielectric, roughness 1, KHR_materials_specular
    factor 0 (src/pathtrace.rs:2810-2817)."""
    return A.material(tuple(rgb) + (1.0,), metallic=0.0, roughness=1.0, specular=0.0, **kw)


def pcg_hash(x):
    x = np.uint32(x)
    with np.errstate(over="ignore"):
        state = np.uint32(x * np.uint32(747796405) + np.uint32(2891336453))
        word = np.uint32(((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737))
    return np.uint32((word >> np.uint32(22)) ^ word)


def pcg_array(seed, n):
    """n chained pcg_hash values from `seed` (uint32), as floats in [0,1)."""
    out = np.empty(n, dtype=np.uint32)
    x = np.uint32(seed)
    for i in range(n):
        x = pcg_hash(x)
        out[i] = x
    return out


def pcg_uniform(seed, n):
    """Vectorised: u_i = to_unit(pcg_hash(pcg_hash(seed) ^ i))."""
    i = np.arange(n, dtype=np.uint32)
    s = pcg_hash(seed)
    with np.errstate(over="ignore"):
        x = i ^ s
        state = (x * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        word = (((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)).astype(np.uint32)
        h = ((word >> np.uint32(22)) ^ word).astype(np.uint32)
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


# ---------------------------------------------------------------------------
# Benchmark workloads.
# ---------------------------------------------------------------------------

def cornell():
    """Cornell box: the five walls of the unit cube [0,1]^3 (10 tris, open
    toward +Z), left red (0.63, 0.065, 0.05), right green (0.14, 0.45,
    0.09), floor/ceiling/back grey 0.73; a short and a tall box rotated
    -17 / +17 degrees about Y (12 tris each, grey 0.73); a 0.25 x 0.25
    emissive quad 1 mm below the ceiling facing -Y, radiance (17, 12, 4),
    single-sided; all Lambert (specular 0). Pinhole camera, yfov 40
    degrees, on the box axis. 36 triangles."""
    s = Scene("cornell")
    white = s.add_material(lambert((0.73, 0.73, 0.73)))
    red = s.add_material(lambert((0.63, 0.065, 0.05)))
    green = s.add_material(lambert((0.14, 0.45, 0.09)))
    light = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(17.0, 12.0, 4.0)))
    walls = [
        ([[0, 0, 0], [0, 0, 1], [1, 0, 1], [1, 0, 0]], white),   # floor, faces +Y
        ([[0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]], white),   # ceiling, faces -Y
        ([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], white),   # back, faces +Z
        ([[0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1]], red),     # left, faces +X
        ([[1, 0, 0], [1, 0, 1], [1, 1, 1], [1, 1, 0]], green),   # right, faces -X
    ]
    for q, m in walls:
        t, _ = quad(q)
        s.add_tris(t, m)
    tall, _ = cuboid((0.0, 0.0, 0.0), (0.3, 0.6, 0.3))
    tall = rotate_y(translate(tall, (0.18, 0.0, 0.2)), 17.0, (0.33, 0.0, 0.35))
    s.add_tris(tall, white)
    short, _ = cuboid((0.0, 0.0, 0.0), (0.3, 0.3, 0.3))
    short = rotate_y(translate(short, (0.52, 0.0, 0.5)), -17.0, (0.67, 0.0, 0.65))
    s.add_tris(short, white)
    h = 0.125
    lq, _ = quad([[0.5 - h, 0.999, 0.5 - h], [0.5 + h, 0.999, 0.5 - h],
                  [0.5 + h, 0.999, 0.5 + h], [0.5 - h, 0.999, 0.5 + h]])
    s.add_tris(lq, light)
    dist = 0.5 / math.tan(math.radians(20.0)) + 1.0
    s.camera = CameraSpec((0.5, 0.5, dist), (0.5, 0.5, 0.0), yfov_deg=40.0)
    s.ambient = 0.0
    return s


def material_sweep(segments=64, rings=32, floor_half=10):
    """7x7 UV spheres (radius 0.4, 64x32 segments, the reference builder's
    4096 triangles each including 128 zero-area pole triangles) on a 1 m
    grid; row r has metallic r/6, column c roughness 0.05 + 0.95 c/6; base
    colour (0.9, 0.6, 0.3), default KHR specular. Checker floor of 1 m
    cells (albedo 0.8 / 0.2, dielectric, roughness 1) as per-cell
    geometry, 20 m x 20 m. One point light (500 cd) and one spot light
    (500 cd, outer half-angle 30 deg, inner 20 deg); uniform sky radiance
    0.3 (RenderOptions::ambient)."""
    s = Scene("material_sweep")
    sp, sn = uv_sphere(0.4, segments, rings)
    for r in range(7):
        for c in range(7):
            m = s.add_material(A.material((0.9, 0.6, 0.3, 1.0), metallic=r / 6.0,
                                          roughness=0.05 + 0.95 * c / 6.0))
            t = (float(c - 3), 0.4, float(r - 3))
            s.add_tris(translate(sp, t), m, normals=sn)
    light_sq = s.add_material(A.material((0.8, 0.8, 0.8, 1.0), metallic=0.0, roughness=1.0))
    dark_sq = s.add_material(A.material((0.2, 0.2, 0.2, 1.0), metallic=0.0, roughness=1.0))
    cells = []
    mats = []
    for i in range(-floor_half, floor_half):
        for j in range(-floor_half, floor_half):
            t, _ = quad([[i, 0, j + 1], [i + 1, 0, j + 1], [i + 1, 0, j], [i, 0, j]])
            cells.append(t)
            m = light_sq if (i + j) % 2 == 0 else dark_sq
            mats += [m, m]
    s.add_tris(np.concatenate(cells), np.asarray(mats, np.uint32))
    s.lights.append(A.light(A.OXR_LIGHT_POINT, intensity=500.0, position=(2.0, 6.0, 5.0)))
    d = _normalize(np.array([3.0, -5.0, 3.0], F32))
    s.lights.append(A.light(A.OXR_LIGHT_SPOT, intensity=500.0, position=(-3.0, 5.0, -3.0),
                            direction=tuple(d.tolist()), inner=math.radians(20.0), outer=math.radians(30.0)))
    s.camera = CameraSpec((0.0, 5.0, 9.0), (0.0, 0.3, 0.0), yfov_deg=40.0)
    s.ambient = 0.3
    return s


def glossy_caustic(segments=64, rings=64):
    """A 10 m x 10 m x 5 m room (12 tris, inward faces, diffuse 0.7) with
    a GGX floor (dielectric, roughness 0.1, base 0.8); a 1 m glass sphere
    (transmission 1, IOR 1.5, volume thickness 1, 64x64 segments = 8192
    tris) at 1.2 m; a 0.1 m emissive quad (radiance 1000) 5 cm above the
    sphere facing down, so most of its light reaches the floor through the
    glass."""
    s = Scene("glossy_caustic")
    wall = s.add_material(lambert((0.7, 0.7, 0.7)))
    floor = s.add_material(A.material((0.8, 0.8, 0.8, 1.0), metallic=0.0, roughness=0.1))
    glass = s.add_material(A.material((1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.0,
                                      transmission=1.0, thickness=1.0, ior=1.5))
    emit = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(1000.0, 1000.0, 1000.0)))
    room, _ = cuboid((-5.0, 0.0, -5.0), (5.0, 5.0, 5.0))
    room = room.reshape(-1, 3, 3)[:, ::-1, :].reshape(-1, 9).copy()  # inward faces
    fl = room[10:12]  # the -Y face (floor)
    s.add_tris(room[:10], wall)
    s.add_tris(fl, floor)
    sp, sn = uv_sphere(0.5, segments, rings)
    s.add_tris(translate(sp, (0.0, 1.2, 0.0)), glass, normals=sn)
    h = 0.05
    y = 1.2 + 0.5 + 0.05
    # Front (emitting) face down, into the sphere: cross(p1 - p0, p3 - p0) = -Y.
    lq, _ = quad([[-h, y, -h], [h, y, -h], [h, y, h], [-h, y, h]])
    s.add_tris(lq, emit)
    s.camera = CameraSpec((0.0, 2.6, 4.6), (0.0, 0.6, 0.0), yfov_deg=40.0)
    s.ambient = 0.0
    return s


def triangle_soup(seed, n_tris, n_emit=4):
    """Seeded random soup: n_tris triangles with corners uniform in the unit
    cube (edge length scaled so the soup stays sparse), random Lambert / GGX
    materials, and n_emit emissive triangles."""
    s = Scene("soup")
    u = pcg_uniform(seed, n_tris * 9 + n_tris * 2 + n_emit * 9).astype(np.float64)
    k = 0
    centers = u[k:k + n_tris * 3].reshape(n_tris, 3)
    k += n_tris * 3
    off = (u[k:k + n_tris * 6].reshape(n_tris, 6) - 0.5) * (0.6 / max(n_tris, 1) ** (1.0 / 3.0))
    k += n_tris * 6
    pos = np.empty((n_tris, 9))
    pos[:, 0:3] = centers
    pos[:, 3:6] = centers + off[:, 0:3]
    pos[:, 6:9] = centers + off[:, 3:6]
    mats = []
    for i in range(8):
        if i % 2 == 0:
            mats.append(s.add_material(lambert((0.3 + 0.08 * i, 0.5, 0.7 - 0.05 * i))))
        else:
            mats.append(s.add_material(A.material((0.8, 0.7, 0.6, 1.0), metallic=0.0, roughness=0.1 + 0.1 * i)))
    mid = (u[k:k + n_tris] * 8).astype(np.uint32).clip(0, 7)
    k += n_tris
    s.add_tris(pos.astype(F32), np.asarray(mats, np.uint32)[mid])
    em = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(20.0, 20.0, 20.0), double_sided=True))
    ep = u[k:k + n_emit * 9].reshape(n_emit, 9) * 0.2 + 0.4
    s.add_tris(ep.astype(F32), em)
    s.camera = CameraSpec((0.5, 0.5, 2.6), (0.5, 0.5, 0.5), yfov_deg=40.0)
    s.ambient = 0.0
    return s


# ---------------------------------------------------------------------------
# The reference's own integrator test scenes (src/pathtrace.rs:3346-3750).
# ---------------------------------------------------------------------------

def furnace(rho=0.5, le=1.0):
    """Closed box of double-sided emissive Lambert walls, viewed from inside
    (src/pathtrace.rs:3348-3359): L = Le / (1 - rho)."""
    s = Scene("furnace")
    m = s.add_material(lambert((rho, rho, rho), emissive=(le, le, le), double_sided=True))
    t, n = cuboid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    s.add_tris(t, m, normals=n)
    s.camera = CameraSpec((0.2, 0.1, 0.4), (0.0, -0.3, -1.0), yfov_deg=math.degrees(1.2))
    s.ambient = 0.0
    return s


def area_light_floor(roughness, metallic):
    """White floor under a down-facing emissive quad
    (src/pathtrace.rs:3378-3412)."""
    s = Scene("area_light")
    fm = A.material((0.8, 0.8, 0.8, 1.0), metallic=metallic, roughness=roughness)
    f = s.add_material(fm)
    lm = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(4.0, 4.0, 4.0)))
    t, n = quad([[-2, 0, 2], [2, 0, 2], [2, 0, -2], [-2, 0, -2]])
    s.add_tris(t, f, normals=n)
    t, n = quad([[-0.5, 1, -0.5], [0.5, 1, -0.5], [0.5, 1, 0.5], [-0.5, 1, 0.5]])
    s.add_tris(t, lm, normals=n)
    s.camera = CameraSpec((0.0, 2.5, 3.0), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.9))
    s.ambient = 0.0
    return s


def glass_cube(volume):
    """Emissive backdrop (Le = 1) seen through a glass cube
    (src/pathtrace.rs:3670-3705)."""
    s = Scene("glass")
    kw = dict(transmission=1.0)
    if volume:
        kw.update(thickness=1.0, attenuation_distance=1.0, attenuation_color=(0.5, 0.5, 0.5))
    g = s.add_material(A.material((1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.0, **kw))
    e = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(1.0, 1.0, 1.0)))
    t, n = cuboid((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
    s.add_tris(t, g, normals=n)
    t, n = quad([[-3, -3, -2], [3, -3, -2], [3, 3, -2], [-3, 3, -2]])
    s.add_tris(t, e, normals=n)
    s.camera = CameraSpec((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.15))
    s.ambient = 0.0
    return s


def _srgb_byte(x):
    x = np.clip(np.asarray(x, np.float64), 0.0, 1.0)
    s = np.where(x <= 0.0031308, x * 12.92, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)
    return np.round(s * 255.0).astype(np.uint8)


def textured_showcase(segments=48, rings=24, parts=None):
    """Every texture path of the tracer in one small scene:
    - a 4 m floor with an RGBA8 checker base-colour texture (0.8 / 0.2 in
      1 m cells through a KHR_texture_transform scale of 4) and a
      metallic-roughness texture;
    - a normal-mapped anisotropic metal sphere (tangents, bump map, an
      anisotropy direction texture) and a clearcoat / sheen sphere with
      extension textures;
    - a double-sided quad emitter with an sRGB emissive gradient;
    - an alpha-MASK quad (a disc texture, cutoff 0.5) and a BLEND quad
      whose alpha comes from vertex colours;
    - a vertex-coloured box, a point light.
    parts: the subset of {"floor", "metal", "coat", "emitter", "mask",
    "blend", "box"} to build (None: all)."""
    s = Scene("textured_showcase")
    want = (lambda k: True) if parts is None else (lambda k: k in parts)
    # Textures.
    n = 64
    yy, xx = np.mgrid[0:n, 0:n]
    cell = ((xx * 4 // n) + (yy * 4 // n)) % 2
    check = np.zeros((n, n, 4), np.uint8)
    check[..., :3] = np.where(cell[..., None] == 0, _srgb_byte(0.8), _srgb_byte(0.2))
    check[..., 3] = 255
    t_check = s.add_texture(check)
    mr = np.zeros((16, 16, 4), np.uint8)
    mr[..., 1] = np.linspace(60, 255, 16, dtype=np.float32)[None, :].astype(np.uint8)  # roughness ramp
    mr[..., 2] = 0
    mr[..., 3] = 255
    t_mr = s.add_texture(mr, min_filter=A.MIN_LINEAR_MIP_NEAREST)
    bump = np.zeros((32, 64, 4), np.float32)
    by, bx = np.mgrid[0:32, 0:64]
    bump[..., 0] = 0.5 + 0.25 * np.sin(2 * np.pi * bx / 8.0)
    bump[..., 1] = 0.5 + 0.25 * np.cos(2 * np.pi * by / 8.0)
    bump[..., 2] = 0.9
    bump[..., 3] = 1.0
    t_bump = s.add_texture(bump, wrap_t=A.WRAP_CLAMP_TO_EDGE)
    aniso = np.zeros((8, 8, 4), np.uint8)
    aniso[..., 0] = 255
    aniso[..., 1] = 128
    aniso[..., 2] = np.linspace(64, 255, 8, dtype=np.float32)[:, None].astype(np.uint8)
    aniso[..., 3] = 255
    t_aniso = s.add_texture(aniso, mag=A.FILTER_NEAREST, min_filter=A.MIN_NEAREST_MIP_LINEAR)
    grad = np.zeros((1, 16, 4), np.uint8)
    grad[0, :, 0] = 255
    grad[0, :, 1] = np.linspace(40, 255, 16).astype(np.uint8)
    grad[0, :, 2] = np.linspace(255, 40, 16).astype(np.uint8)
    grad[0, :, 3] = 255
    t_grad = s.add_texture(grad, wrap_s=A.WRAP_MIRRORED_REPEAT)
    disc = np.zeros((32, 32, 4), np.uint8)
    dy, dx = np.mgrid[0:32, 0:32]
    disc[..., :3] = (200, 120, 60)
    disc[..., 3] = np.where((dx - 15.5) ** 2 + (dy - 15.5) ** 2 < 14.0 ** 2, 255, 0).astype(np.uint8)
    t_disc = s.add_texture(disc)
    cc = np.zeros((8, 8, 4), np.uint8)
    cc[..., 0] = np.linspace(80, 255, 8).astype(np.uint8)[None, :]
    cc[..., 1] = np.linspace(255, 40, 8).astype(np.uint8)[:, None]
    cc[..., 2] = 200
    cc[..., 3] = 180
    t_cc = s.add_texture(cc)
    # Floor.
    floor = s.add_material(A.material((1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.8))
    t, nr, uv = quad_uv([[-2, 0, 2], [2, 0, 2], [2, 0, -2], [-2, 0, -2]])
    if want("floor"):
        s.add_tris(t, floor, normals=nr, uv0=uv)
    s.set_material_textures(floor, base_color=dict(texture=t_check, scale=(4.0, 4.0)), metallic_roughness=t_mr)
    # Normal-mapped anisotropic sphere.
    metal = s.add_material(A.material((0.9, 0.7, 0.4, 1.0), metallic=1.0, roughness=0.35, anisotropy=0.6,
                                      anisotropy_rotation=0.3))
    sp, sn, suv, stg = uv_sphere_textured(0.5, segments, rings, center=(-0.7, 0.5, 0.0))
    if want("metal"):
        s.add_tris(sp, metal, normals=sn, uv0=suv, tangents=stg)
    s.set_material_textures(metal, normal=t_bump, anisotropy=t_aniso, normal_scale=0.8)
    # Clearcoat / sheen sphere with extension textures (uv on TEXCOORD_1).
    coat = s.add_material(A.material((0.3, 0.5, 0.8, 1.0), metallic=0.0, roughness=0.6, clearcoat=1.0,
                                     clearcoat_roughness=0.2, sheen_color=(0.6, 0.6, 0.6), sheen_roughness=0.5))
    cp, cn, cuv, ctg = uv_sphere_textured(0.4, segments, rings, center=(0.6, 0.4, 0.3))
    if want("coat"):
        s.add_tris(cp, coat, normals=cn, uv1=cuv, tangents=ctg)
    s.set_material_textures(coat, clearcoat=dict(texture=t_cc, uv_set=1), clearcoat_normal=dict(texture=t_bump, uv_set=1),
                            sheen_color=dict(texture=t_cc, uv_set=1), sheen_roughness=dict(texture=t_cc, uv_set=1),
                            clearcoat_normal_scale=0.5)
    # Textured emitter (double-sided).
    em = s.add_material(lambert((0.0, 0.0, 0.0), emissive=(6.0, 6.0, 6.0), double_sided=True))
    t, nr, uv = quad_uv([[-0.5, 2.2, -0.5], [0.5, 2.2, -0.5], [0.5, 2.2, 0.5], [-0.5, 2.2, 0.5]],
                        uv=((0, 0), (2, 0), (2, 1), (0, 1)))
    if want("emitter"):
        s.add_tris(t, em, normals=nr, uv0=uv)
    s.set_material_textures(em, emissive=t_grad)
    # Alpha-masked disc quad and a vertex-colour BLEND quad.
    mask = A.material((1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.7, double_sided=True)
    mask.alpha_mode = A.OXR_ALPHA_MASK
    mask.alpha_cutoff = 0.5
    mk = s.add_material(mask)
    t, nr, uv = quad_uv([[1.2, 0.05, -1.0], [1.2, 0.05, 0.2], [1.2, 1.25, 0.2], [1.2, 1.25, -1.0]])
    if want("mask"):
        s.add_tris(t, mk, normals=nr, uv0=uv)
    s.set_material_textures(mk, base_color=t_disc)
    blend = A.material((0.9, 0.9, 0.9, 1.0), metallic=0.0, roughness=0.5, double_sided=True)
    blend.alpha_mode = A.OXR_ALPHA_BLEND
    bl = s.add_material(blend)
    t, nr = quad([[-1.6, 0.05, 0.6], [-0.4, 0.05, 0.6], [-0.4, 1.0, 0.6], [-1.6, 1.0, 0.6]])
    cols = np.tile(np.array([1.0, 0.6, 0.6, 0.9, 0.6, 1.0, 0.6, 0.2, 0.6, 0.6, 1.0, 0.5], np.float32), (2, 1))
    if want("blend"):
        s.add_tris(t, bl, normals=nr, colors=cols)
    # Vertex-coloured box.
    box = s.add_material(lambert((1.0, 1.0, 1.0)))
    bt, bn = cuboid((0.2, 0.0, -1.4), (0.8, 0.5, -0.8))
    bc = np.tile(np.array([0.9, 0.2, 0.2, 1.0, 0.2, 0.9, 0.2, 1.0, 0.2, 0.2, 0.9, 1.0], np.float32), (len(bt), 1))
    if want("box"):
        s.add_tris(bt, box, colors=bc)
    s.lights.append(A.light(A.OXR_LIGHT_POINT, intensity=6.0, position=(1.5, 2.5, 1.5), direction=(0, -1, 0)))
    s.camera = CameraSpec((0.0, 1.6, 3.4), (0.0, 0.5, -0.2), yfov_deg=45.0)
    s.ambient = 0.05
    return s


def lit_plane(rho=1.0, with_point=True):
    """Lambert plane (ρ) with an optional 4 cd point light 1 m above the
    origin (src/pathtrace.rs:3717-3742)."""
    s = Scene("plane")
    m = s.add_material(lambert((rho, rho, rho)))
    t, n = quad([[-1, 0, 1], [1, 0, 1], [1, 0, -1], [-1, 0, -1]])
    s.add_tris(t, m, normals=n)
    if with_point:
        s.lights.append(A.light(A.OXR_LIGHT_POINT, intensity=4.0, position=(0.0, 1.0, 0.0),
                                direction=(0.0, -1.0, 0.0)))
    s.camera = CameraSpec((0.0, 3.0, 0.001), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.05))
    s.ambient = 0.0
    return s


WORKLOADS = {
    "cornell": cornell,
    "material_sweep": material_sweep,
    "glossy_caustic": glossy_caustic,
}


def reference_box():
    """The reference's `testscenes::cornell_box` (src/testscenes.rs:305-385):
    [-1,1]^3 room open toward +Z, red left / green right / white walls,
    a tall and a short block, a point light (6 cd, colour (1, .95, .85))
    under the ceiling; camera at (0, 0, 3.4), yfov 0.75 rad. Materials are
    `diffuse_material` (default KHR specular)."""
    s = Scene("reference_box")
    white = s.add_material(A.material((0.75, 0.75, 0.75, 1.0), metallic=0.0, roughness=1.0))
    red = s.add_material(A.material((0.65, 0.05, 0.05, 1.0), metallic=0.0, roughness=1.0))
    green = s.add_material(A.material((0.12, 0.45, 0.15, 1.0), metallic=0.0, roughness=1.0))
    walls = [
        ([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]], white),
        ([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], white),
        ([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], white),
        ([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]], red),
        ([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]], green),
    ]
    for q, m in walls:
        t, n = quad(q)
        s.add_tris(t, m, normals=n)
    t, n = cuboid((-0.6, -1.0, -0.6), (-0.1, 0.2, -0.1))
    s.add_tris(t, white, normals=n)
    t, n = cuboid((0.15, -1.0, -0.1), (0.65, -0.4, 0.4))
    s.add_tris(t, white, normals=n)
    s.lights.append(A.light(A.OXR_LIGHT_POINT, color=(1.0, 0.95, 0.85), intensity=6.0,
                            position=(0.0, 0.85, 0.0), direction=(0.0, -1.0, 0.0)))
    s.camera = CameraSpec((0.0, 0.0, 3.4), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.75))
    s.ambient = 0.0
    return s



def coplanar_overlaps(seed=5, n_base=40, copies=4, n_rotated=16):
    """Triangles that overlap in the plane z = 0, for the equal-t
    tie-break of closest hit (the lowest global id wins). n_base random
    triangles with two-decimal corners in [0, 1]^2, each present `copies`
    times with identical corners (so identical hit t) at randomly permuted
    global ids, then n_rotated copies of base triangles with their vertices
    rotated (same plane and region; t may differ in the last place).
    One grey double-sided material. Returns (scene, corners (n, 3, 2) in
    global-id order, base index per triangle (-1 for the rotated ones))."""
    rng = np.random.default_rng(seed)
    base = np.round(rng.random((n_base, 3, 2)), 2)
    which = np.repeat(np.arange(n_base), copies)
    perm = rng.permutation(len(which))
    xy = base[which[perm]]
    owner = which[perm]
    rot = rng.choice(n_base, n_rotated, replace=False)
    xy = np.concatenate([xy, np.roll(base[rot], 1, axis=1)])
    owner = np.concatenate([owner, np.full(n_rotated, -1)])
    pos = np.zeros((len(xy), 3, 3), F32)
    pos[:, :, :2] = xy
    sc = Scene("coplanar")
    m = sc.add_material(A.material((0.5, 0.5, 0.5, 1.0), double_sided=True))
    sc.add_tris(pos.reshape(-1, 9), m)
    sc.camera = CameraSpec((0.5, 0.5, 2.0), (0.5, 0.5, 0.0), yfov_deg=40.0)
    return sc, xy.astype(np.float64), owner

# ---------------------------------------------------------------------------
# Large-scene throughput: displaced terrain + emitters + instanced spheres.
# ---------------------------------------------------------------------------

def _hash2(ix, iz, seed):
    """pcg-based lattice hash -> [0, 1) (uint32 arithmetic, vectorised)."""
    with np.errstate(over="ignore"):
        x = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iz.astype(np.uint32) * np.uint32(19349663)) ^ np.uint32(seed)
        state = (x * np.uint32(747796405) + np.uint32(2891336453)).astype(np.uint32)
        word = (((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)).astype(np.uint32)
        h = ((word >> np.uint32(22)) ^ word).astype(np.uint32)
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def fbm(x, z, seed, octaves=6, lacunarity=2.0, gain=0.5, base_freq=4.0):
    """Value-noise fBm in [-1, 1]-ish (smoothstep-interpolated lattice)."""
    out = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx), np.floor(fz)
        tx, tz = fx - ix, fz - iz
        tx = tx * tx * (3 - 2 * tx)
        tz = tz * tz * (3 - 2 * tz)
        ix = ix.astype(np.int64)
        iz = iz.astype(np.int64)
        s = seed + 1013 * o
        a = _hash2(ix, iz, s)
        b = _hash2(ix + 1, iz, s)
        c = _hash2(ix, iz + 1, s)
        d = _hash2(ix + 1, iz + 1, s)
        v = (a * (1 - tx) + b * tx) * (1 - tz) + (c * (1 - tx) + d * tx) * tz
        out += amp * (2.0 * v - 1.0)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return out / norm


def icosphere1():
    """Subdivision-1 icosphere: 80 triangles on the unit sphere (CCW out),
    with smooth (radial) corner normals."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    tris = []
    for a, b, c in f:
        A, B, Cc = v[a], v[b], v[c]
        ab = (A + B) / np.linalg.norm(A + B)
        bc = (B + Cc) / np.linalg.norm(B + Cc)
        ca = (Cc + A) / np.linalg.norm(Cc + A)
        tris += [[A, ab, ca], [B, bc, ab], [Cc, ca, bc], [ab, bc, ca]]
    tris = np.array(tris, np.float64)  # (80, 3, 3)
    # Orient CCW outward.
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = (n * tris.mean(axis=1)).sum(1) < 0
    tris[flip] = tris[flip][:, ::-1]
    return tris


def terrain(seed=7, grid=2048, n_emit=10000, n_spheres=50000, amplitude=0.1):
    """Procedural displaced terrain over [0, 1]^2: a grid x grid heightfield
    (2·grid² tris, (grid+1)² vertices) with 6-octave fBm heights of the
    given amplitude; n_emit emissive triangles (double-sided, radiance
    log-uniform 1..100, edge ~4 mm) scattered uniformly in the volume above
    it; n_spheres instanced 80-tri icospheres (radius 2-6 mm, resting on the
    surface) whose materials are drawn 25 % each from thin transmission
    (IOR 1.5), clearcoat, sheen and GGX metal (roughness U(0.05, 0.5)).
    Defaults give 8,388,608 + 10,000 + 4,000,000 = 12,398,608 tris."""
    s = Scene("terrain")
    g = grid
    xs = np.linspace(0.0, 1.0, g + 1)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    H = amplitude * fbm(X, Z, seed)
    P = np.stack([X, H, Z], -1).astype(np.float32)  # (g+1, g+1, 3)
    p00 = P[:-1, :-1].reshape(-1, 3)
    p10 = P[:-1, 1:].reshape(-1, 3)
    p01 = P[1:, :-1].reshape(-1, 3)
    p11 = P[1:, 1:].reshape(-1, 3)
    # CCW seen from +Y: (p00, p01, p11), (p00, p11, p10).
    t1 = np.concatenate([p00, p01, p11], 1)
    t2 = np.concatenate([p00, p11, p10], 1)
    tris = np.empty((2 * len(t1), 9), np.float32)
    tris[0::2] = t1
    tris[1::2] = t2
    ground = s.add_material(A.material((0.45, 0.5, 0.35, 1.0), metallic=0.0, roughness=0.8))
    s.add_tris(tris, ground)
    del tris, t1, t2
    u = pcg_uniform(seed * 7919 + 1, n_emit * 9)
    ctr = np.stack([u[:n_emit], 0.05 + 0.25 * u[n_emit:2 * n_emit], u[2 * n_emit:3 * n_emit]], 1)
    off = (u[3 * n_emit:9 * n_emit].reshape(n_emit, 2, 3) - 0.5) * 0.008
    em = np.concatenate([ctr, ctr + off[:, 0], ctr + off[:, 1]], 1).astype(np.float32)
    ue = pcg_uniform(seed * 7919 + 2, n_emit)
    le = np.exp(np.log(1.0) + ue * (np.log(100.0) - np.log(1.0)))
    # One material per emitter radiance (log-uniform 1..100).
    emat = material_array(n_emit, lambert((0.0, 0.0, 0.0), double_sided=True))
    emat["emissive"] = np.repeat(le[:, None], 3, 1).astype(np.float32)
    mids = (s.add_material_array(emat) + np.arange(n_emit)).astype(np.uint32)
    s.add_tris(em, mids)
    ico = icosphere1()  # (80, 3, 3) unit
    us = pcg_uniform(seed * 7919 + 3, n_spheres * 6)
    cx = us[:n_spheres]
    cz = us[n_spheres:2 * n_spheres]
    r = 0.002 + 0.004 * us[2 * n_spheres:3 * n_spheres]
    cy = amplitude * fbm(cx, cz, seed) + r
    kind = (us[3 * n_spheres:4 * n_spheres] * 4).astype(int).clip(0, 3)
    p1 = us[4 * n_spheres:5 * n_spheres]
    p2 = us[5 * n_spheres:6 * n_spheres]
    sm = material_array(n_spheres, A.material((1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.5))
    t0, t1, t2, t3 = (kind == 0), (kind == 1), (kind == 2), (kind == 3)
    bc = np.ones((n_spheres, 4), np.float32)
    bc[t0, :3] = (0.95, 0.97, 1.0)
    bc[t1, 0], bc[t1, 1], bc[t1, 2] = 0.2 + 0.6 * p1[t1], 0.2, 0.6 * p2[t1]
    bc[t2, 0], bc[t2, 1], bc[t2, 2] = 0.3, 0.1 + 0.3 * p2[t2], 0.1
    bc[t3, 0], bc[t3, 1], bc[t3, 2] = 0.95, 0.64 + 0.3 * p2[t3], 0.54
    sm["base_color"] = bc
    rough = np.full(n_spheres, 0.5, np.float32)
    rough[t0] = 0.05 + 0.2 * p1[t0]
    rough[t1] = 0.6
    rough[t2] = 0.8
    rough[t3] = 0.05 + 0.45 * p1[t3]
    sm["roughness"] = rough
    sm["metallic"] = np.where(t3, 1.0, 0.0).astype(np.float32)
    sm["transmission"] = np.where(t0, 1.0, 0.0).astype(np.float32)
    sm["clearcoat"] = np.where(t1, 1.0, 0.0).astype(np.float32)
    sm["clearcoat_roughness"] = np.where(t1, 0.1, 0.0).astype(np.float32)
    sh = np.zeros((n_spheres, 3), np.float32)
    sh[t2, 0], sh[t2, 1], sh[t2, 2] = 0.9, 0.7 * p1[t2], 0.7
    sm["sheen_color"] = sh
    sm["sheen_roughness"] = np.where(t2, 0.3 + 0.5 * p2, 0.0).astype(np.float32)
    mids = (s.add_material_array(sm) + np.arange(n_spheres)).astype(np.uint32)
    pos = (ico[None] * r[:, None, None, None] + np.stack([cx, cy, cz], 1)[:, None, None, :]).astype(np.float32)
    nrm = np.broadcast_to(ico[None], pos.shape).astype(np.float32)
    s.add_tris(pos.reshape(-1, 9), np.repeat(mids, 80), normals=nrm.reshape(-1, 9))
    s.camera = CameraSpec((0.5, 0.45, 1.25), (0.5, 0.0, 0.45), yfov_deg=45.0)
    s.ambient = 0.0
    return s


WORKLOADS["terrain"] = terrain
