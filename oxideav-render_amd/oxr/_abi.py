"""ctypes mirror of include/oxr/pathtrace.h (the C ABI of liboxr.so).

Field order and types follow the header exactly; `tests/test_abi.py`
checks the sizes against the header's layout.
"""
import ctypes as C

OXR_OK = 0
OXR_ERR_INVALID = -1
OXR_ERR_DEVICE = -2
OXR_ERR_STATE = -3
OXR_ERR_MEMORY = -4

OXR_ALPHA_OPAQUE = 0
OXR_ALPHA_MASK = 1
OXR_ALPHA_BLEND = 2

OXR_MAT_DOUBLE_SIDED = 1
OXR_MAT_UNLIT = 2

OXR_LIGHT_DIRECTIONAL = 0
OXR_LIGHT_POINT = 1
OXR_LIGHT_SPOT = 2

OXR_TRI_HAS_NORMALS = 1

OXR_PROJ_PERSPECTIVE = 0
OXR_PROJ_ORTHOGRAPHIC = 1

OXR_STRATEGY_MIS = 0
OXR_STRATEGY_LIGHT_ONLY = 1
OXR_STRATEGY_BSDF_ONLY = 2

OXR_TONEMAP_CLAMP = 0
OXR_TONEMAP_REINHARD = 1
OXR_TONEMAP_ACES = 2

OXR_HIT_MISS = 0xFFFFFFFF
OXR_HIT_BACK = 0x80000000

f3 = C.c_float * 3
f4 = C.c_float * 4


class Material(C.Structure):
    _fields_ = [
        ("base_color", f4),
        ("metallic", C.c_float),
        ("roughness", C.c_float),
        ("emissive", f3),
        ("flags", C.c_uint32),
        ("alpha_mode", C.c_uint32),
        ("alpha_cutoff", C.c_float),
        ("ior", C.c_float),
        ("specular", C.c_float),
        ("specular_color", f3),
        ("transmission", C.c_float),
        ("thickness", C.c_float),
        ("attenuation_color", f3),
        ("attenuation_distance", C.c_float),
        ("clearcoat", C.c_float),
        ("clearcoat_roughness", C.c_float),
        ("sheen_color", f3),
        ("sheen_roughness", C.c_float),
        ("anisotropy", C.c_float),
        ("anisotropy_rotation", C.c_float),
        ("iridescence", C.c_float),
        ("iridescence_ior", C.c_float),
        ("iridescence_thickness", C.c_float),
        ("diffuse_transmission", C.c_float),
        ("diffuse_transmission_color", f3),
        ("dispersion", C.c_float),
    ]


def material(base=(1.0, 1.0, 1.0, 1.0), metallic=1.0, roughness=1.0, **kw):
    """A material with the glTF defaults (white, metallic 1, roughness 1,
    IOR 1.5, core specular) and keyword overrides."""
    m = Material()
    b = list(base) + [1.0] * (4 - len(base))
    m.base_color = f4(*b)
    m.metallic = metallic
    m.roughness = roughness
    m.emissive = f3(0.0, 0.0, 0.0)
    m.flags = 0
    m.alpha_mode = OXR_ALPHA_OPAQUE
    m.alpha_cutoff = 0.5
    m.ior = 1.5
    m.specular = 1.0
    m.specular_color = f3(1.0, 1.0, 1.0)
    m.transmission = 0.0
    m.thickness = 0.0
    m.attenuation_color = f3(1.0, 1.0, 1.0)
    m.attenuation_distance = float("inf")
    m.clearcoat = 0.0
    m.clearcoat_roughness = 0.0
    m.sheen_color = f3(0.0, 0.0, 0.0)
    m.sheen_roughness = 0.0
    m.anisotropy = 0.0
    m.anisotropy_rotation = 0.0
    m.iridescence = 0.0
    m.iridescence_ior = 1.3
    m.iridescence_thickness = 400.0
    m.diffuse_transmission = 0.0
    m.diffuse_transmission_color = f3(1.0, 1.0, 1.0)
    m.dispersion = 0.0
    for k, v in kw.items():
        if k == "double_sided":
            m.flags = (m.flags | OXR_MAT_DOUBLE_SIDED) if v else (m.flags & ~OXR_MAT_DOUBLE_SIDED)
        elif k == "unlit":
            m.flags = (m.flags | OXR_MAT_UNLIT) if v else (m.flags & ~OXR_MAT_UNLIT)
        elif isinstance(getattr(m, k), C.Array):
            setattr(m, k, type(getattr(m, k))(*v))
        else:
            setattr(m, k, v)
    return m


class Light(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("color", f3),
        ("intensity", C.c_float),
        ("position", f3),
        ("direction", f3),
        ("range", C.c_float),
        ("inner_cone_angle", C.c_float),
        ("outer_cone_angle", C.c_float),
    ]


def light(kind, color=(1.0, 1.0, 1.0), intensity=1.0, position=(0.0, 0.0, 0.0),
          direction=(0.0, 0.0, -1.0), range_=0.0, inner=0.0, outer=0.7853982):
    l = Light()
    l.kind = kind
    l.color = f3(*color)
    l.intensity = intensity
    l.position = f3(*position)
    l.direction = f3(*direction)
    l.range = range_
    l.inner_cone_angle = inner
    l.outer_cone_angle = outer
    return l


TEXFMT_RGBA8, TEXFMT_RGBA32F = 0, 1
FILTER_NEAREST, FILTER_LINEAR = 0, 1
(MIN_NEAREST, MIN_LINEAR, MIN_NEAREST_MIP_NEAREST, MIN_LINEAR_MIP_NEAREST, MIN_NEAREST_MIP_LINEAR,
 MIN_LINEAR_MIP_LINEAR) = range(6)
WRAP_REPEAT, WRAP_CLAMP_TO_EDGE, WRAP_MIRRORED_REPEAT = 0, 1, 2
TRI_HAS_NORMALS, TRI_HAS_UV0, TRI_HAS_UV1, TRI_HAS_TANGENTS, TRI_HAS_COLORS = 1, 2, 4, 8, 16


class Texture(C.Structure):
    """oxr_texture."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32), ("pixels", C.c_void_p),
                ("mag_filter", C.c_uint32), ("min_filter", C.c_uint32), ("wrap_s", C.c_uint32),
                ("wrap_t", C.c_uint32)]


class TextureBinding(C.Structure):
    """oxr_texture_binding."""
    _fields_ = [("texture", C.c_int32), ("uv_set", C.c_uint32), ("has_transform", C.c_uint32),
                ("offset", C.c_float * 2), ("rotation", C.c_float), ("scale", C.c_float * 2)]


TEXTURE_SLOTS = ("base_color", "metallic_roughness", "normal", "emissive", "specular", "specular_color",
                 "transmission", "thickness", "clearcoat", "clearcoat_roughness", "clearcoat_normal", "sheen_color",
                 "sheen_roughness", "anisotropy", "iridescence", "iridescence_thickness", "diffuse_transmission",
                 "diffuse_transmission_color")


class MaterialTextures(C.Structure):
    """oxr_material_textures."""
    _fields_ = [(n, TextureBinding) for n in TEXTURE_SLOTS] + [
        ("normal_scale", C.c_float), ("clearcoat_normal_scale", C.c_float),
        ("iridescence_thickness_min", C.c_float), ("pad", C.c_float)]


def no_binding():
    b = TextureBinding()
    b.texture = -1
    b.scale[0] = b.scale[1] = 1.0
    return b


def material_textures():
    """An oxr_material_textures with every slot empty and the glTF
    default scales."""
    m = MaterialTextures()
    for n in TEXTURE_SLOTS:
        setattr(m, n, no_binding())
    m.normal_scale = 1.0
    m.clearcoat_normal_scale = 1.0
    m.iridescence_thickness_min = 100.0
    return m


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_tris", C.c_uint32),
        ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)),
        ("tri_flags", C.POINTER(C.c_uint32)),
        ("material_ids", C.POINTER(C.c_uint32)),
        ("num_materials", C.c_uint32),
        ("materials", C.POINTER(Material)),
        ("num_lights", C.c_uint32),
        ("lights", C.POINTER(Light)),
        ("num_textures", C.c_uint32),
        ("textures", C.POINTER(Texture)),
        ("material_textures", C.POINTER(MaterialTextures)),
        ("uv0", C.POINTER(C.c_float)),
        ("uv1", C.POINTER(C.c_float)),
        ("tangents", C.POINTER(C.c_float)),
        ("colors", C.POINTER(C.c_float)),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("eye", f3),
        ("forward", f3),
        ("side", f3),
        ("up", f3),
        ("half_w", C.c_float),
        ("half_h", C.c_float),
        ("projection", C.c_uint32),
    ]


class PtOptions(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("samples_per_pixel", C.c_uint32),
        ("max_bounces", C.c_uint32),
        ("rr_start", C.c_uint32),
        ("clamp", C.c_float),
        ("seed", C.c_uint32),
        ("strategy", C.c_uint32),
        ("ambient", C.c_float),
    ]


class Display(C.Structure):
    _fields_ = [
        ("background", C.c_uint8 * 4),
        ("tone_map", C.c_uint32),
        ("exposure", C.c_float),
    ]


class Ray(C.Structure):
    _fields_ = [("o", f3), ("t_min", C.c_float), ("d", f3), ("t_max", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_uint32), ("u", C.c_float), ("v", C.c_float)]


class PtStats(C.Structure):
    _fields_ = [
        ("paths", C.c_uint64),
        ("closest_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("ms_closest", C.c_double),
        ("ms_anyhit", C.c_double),
        ("ms_shade", C.c_double),
        ("ms_other", C.c_double),
        ("launches_closest", C.c_uint32),
        ("launches_anyhit", C.c_uint32),
        ("graph_replays", C.c_uint32),
        ("graph_captures", C.c_uint32),
        ("closest_nodes", C.c_uint64),
        ("closest_tris", C.c_uint64),
        ("anyhit_nodes", C.c_uint64),
        ("anyhit_tris", C.c_uint64),
    ]


class TuneResult(C.Structure):
    """oxr_tune_result."""
    _fields_ = [("path_budget", C.c_uint32), ("shade_waves", C.c_uint32), ("trials", C.c_uint32),
                ("xcd_segments", C.c_uint32), ("paths_per_s", C.c_double), ("worst_paths_per_s", C.c_double),
                ("refill_closest", C.c_uint32), ("refill_any", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_uint32),
        ("num_leaves", C.c_uint32),
        ("num_tris", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("sah_cost", C.c_double),
        ("build_ms", C.c_double),
    ]


class JobStatus(C.Structure):
    """oxr_job_status."""
    _fields_ = [("state", C.c_uint32), ("samples", C.c_uint32), ("target", C.c_uint32), ("rounds", C.c_uint32),
                ("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_done_ms", C.c_double)]


class SrvStats(C.Structure):
    """oxr_srv_stats."""
    _fields_ = [("rounds", C.c_uint64), ("paths", C.c_uint64), ("jobs_done", C.c_uint64), ("active", C.c_uint32),
                ("queued", C.c_uint32), ("bytes_resident", C.c_uint64), ("busy_ms", C.c_double),
                ("graph_captures", C.c_uint32), ("graph_replays", C.c_uint32), ("single_scene_rounds", C.c_uint64),
                ("scenes", C.c_uint32), ("pad", C.c_uint32)]


OXR_BUILD_FAST = 1

OXR_JOB_QUEUED = 0
OXR_JOB_ACTIVE = 1
OXR_JOB_DONE = 2


GUIDE_LEVELS = 16
GUIDE_FEATURES = 2
GUIDE_HIDDEN = 64
GUIDE_OUTPUTS = 48
GUIDE_LOBES = 8
GUIDE_MLP_PARAMS = 64 * 32 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 48 * 64 + 48


class GuideConfig(C.Structure):
    """oxr_guide_config."""
    _fields_ = [("log2_table", C.c_uint32), ("base_res", C.c_uint32), ("max_res", C.c_uint32), ("seed", C.c_uint32),
                ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("mix", C.c_float), ("batch", C.c_uint32),
                ("steps", C.c_uint32), ("record_shift", C.c_uint32), ("max_steps", C.c_uint32),
                ("bounces", C.c_uint32)]


class GuideStats(C.Structure):
    """oxr_guide_stats."""
    _fields_ = [("steps", C.c_uint64), ("samples", C.c_uint64), ("collected", C.c_uint64), ("last_loss", C.c_float),
                ("pad", C.c_float), ("train_ms", C.c_double), ("infer_ms", C.c_double)]


def guide_config(log2_table=19, base_res=16, max_res=2048, seed=0, aabb_min=(0.0, 0.0, 0.0), aabb_max=(0.0, 0.0, 0.0),
                 lr=5e-3, beta1=0.9, beta2=0.99, eps=1e-15, mix=0.5, batch=65536, steps=4, record_shift=3,
                 max_steps=0, bounces=1):
    """oxr_guide_default_config with overrides."""
    c = GuideConfig()
    c.log2_table, c.base_res, c.max_res, c.seed = log2_table, base_res, max_res, seed
    c.aabb_min = f3(*aabb_min)
    c.aabb_max = f3(*aabb_max)
    c.lr, c.beta1, c.beta2, c.eps, c.mix = lr, beta1, beta2, eps, mix
    c.batch, c.steps, c.record_shift, c.max_steps, c.bounces = batch, steps, record_shift, max_steps, bounces
    return c


class MaterialSample(C.Structure):
    _fields_ = [
        ("base_color", f3),
        ("metallic", C.c_float),
        ("roughness", C.c_float),
        ("normal", f3),
        ("ior", C.c_float),
        ("specular", C.c_float),
        ("specular_color", f3),
        ("transmission", C.c_float),
        ("thickness", C.c_float),
        ("clearcoat", C.c_float),
        ("clearcoat_roughness", C.c_float),
        ("clearcoat_normal", f3),
        ("sheen_color", f3),
        ("sheen_roughness", C.c_float),
        ("anisotropy", C.c_float),
        ("anisotropy_direction", f3),
        ("iridescence", C.c_float),
        ("iridescence_ior", C.c_float),
        ("iridescence_thickness", C.c_float),
        ("diffuse_transmission", C.c_float),
        ("diffuse_transmission_color", f3),
    ]


def material_sample(base=(1.0, 1.0, 1.0), metallic=1.0, roughness=1.0, **kw):
    """MaterialSample::default() (src/trace.rs:269-305) with overrides."""
    m = MaterialSample()
    m.base_color = f3(*base)
    m.metallic = metallic
    m.roughness = roughness
    m.normal = f3(0.0, 0.0, 1.0)
    m.ior = 1.5
    m.specular = 1.0
    m.specular_color = f3(1.0, 1.0, 1.0)
    m.transmission = 0.0
    m.thickness = 0.0
    m.clearcoat = 0.0
    m.clearcoat_roughness = 0.0
    m.clearcoat_normal = f3(0.0, 0.0, 1.0)
    m.sheen_color = f3(0.0, 0.0, 0.0)
    m.sheen_roughness = 0.0
    m.anisotropy = 0.0
    m.anisotropy_direction = f3(0.0, 0.0, 0.0)
    m.iridescence = 0.0
    m.iridescence_ior = 1.3
    m.iridescence_thickness = 400.0
    m.diffuse_transmission = 0.0
    m.diffuse_transmission_color = f3(1.0, 1.0, 1.0)
    for k, v in kw.items():
        if isinstance(getattr(m, k), C.Array):
            setattr(m, k, f3(*v))
        else:
            setattr(m, k, v)
    return m
