"""The oracle's texture restatement (src/texture.rs:225-515,
src/trace.rs:492-606, 640-895, src/pathtrace.rs:1484-1500, 1951-1966),
pinned by the reference's own texture tests (src/texture.rs:1060-1148:
nearest and wrap modes, the bilinear midpoint, the sRGB decode, the mip box
filter, non-power-of-two chains, gradient level selection) and by closed
forms for what those leave open (ray-cone level, normal mapping, the
emissive power of a textured emitter). The KHR_texture_transform matrix
is oxideav-mesh3d's (crates.io, not vendored); it is checked against the
extension's own definition, and stays parity unpinned against the
reference."""
import math

import numpy as np
import pytest

import oracle as O
from oxr import _abi as A
from oxr import scenes as S


def quad_scene(tex_pixels=None, sampler=None, uv=None, tangents=None, colors=None, slots=None, mat=None,
               normal_scale=1.0, extra=None):
    """A unit quad in z = 0 facing +z (two triangles) with uv = xy."""
    s = S.Scene("texquad")
    m = s.add_material(mat if mat is not None else S.lambert((1.0, 1.0, 1.0)))
    p = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    pos = np.stack([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])])
    if uv is None:
        uv = np.stack([np.concatenate([p[0, :2], p[1, :2], p[2, :2]]), np.concatenate([p[0, :2], p[2, :2], p[3, :2]])])
    s.add_tris(pos, m, uv0=uv, tangents=tangents, colors=colors)
    if tex_pixels is not None:
        t = s.add_texture(tex_pixels, **(sampler or {}))
        s.set_material_textures(m, normal_scale=normal_scale, **(slots or {"base_color": t}))
    if extra:
        extra(s)
    s.camera = S.CameraSpec((0.5, 0.5, 2.0), (0.5, 0.5, 0.0), yfov_deg=40.0)
    return s


def two_by_one():
    return np.array([[[0, 0, 0, 255], [255, 255, 255, 255]]], np.uint8)


NEAREST = dict(mag=A.FILTER_NEAREST, min_filter=A.MIN_NEAREST)


def test_nearest_and_wrap_modes():
    """src/texture.rs:1060-1098."""
    def tex(ws, wt):
        sc = quad_scene(two_by_one(), dict(NEAREST, wrap_s=ws, wrap_t=wt))
        return O.OracleScene(sc)
    rep = tex(A.WRAP_REPEAT, A.WRAP_REPEAT)
    s = lambda o, u: float(o.tex_sample(0, [[u, 0.5]], 0.0)[0, 0])  # noqa: E731
    assert s(rep, 0.25) == 0.0 and s(rep, 0.75) == 1.0 and s(rep, 1.25) == 0.0
    clamp = tex(A.WRAP_CLAMP_TO_EDGE, A.WRAP_CLAMP_TO_EDGE)
    assert s(clamp, 1.25) == 1.0 and s(clamp, -3.0) == 0.0
    mirror = tex(A.WRAP_MIRRORED_REPEAT, A.WRAP_REPEAT)
    assert s(mirror, 1.25) == 1.0 and s(mirror, 1.75) == 0.0


def test_bilinear_midpoint_and_srgb_decode():
    """src/texture.rs:1100-1116 (default sampler: linear, trilinear)."""
    o = O.OracleScene(quad_scene(two_by_one(), dict(wrap_s=A.WRAP_CLAMP_TO_EDGE, wrap_t=A.WRAP_CLAMP_TO_EDGE)))
    assert abs(o.tex_sample(0, [[0.5, 0.5]])[0, 0] - 0.5) < 1e-6
    g = O.OracleScene(quad_scene(np.full((1, 1, 4), 128, np.uint8)))
    assert abs(g.tex_sample(0, [[0.3, 0.3]], srgb=True)[0, 0] - 0.2158) < 1e-3
    assert abs(g.tex_sample(0, [[0.3, 0.3]], srgb=False)[0, 0] - 128.0 / 255.0) < 1e-6
    assert g.tex_sample(0, [[0.3, 0.3]], srgb=True)[0, 3] == pytest.approx(128.0 / 255.0)  # alpha stays linear


def checker4():
    px = np.zeros((4, 4, 4), np.uint8)
    for y in range(4):
        for x in range(4):
            v = 255 if (x + y) % 2 == 0 else 0
            px[y, x] = (v, v, v, 255)
    return px


def test_mip_chain_box_filters_to_mean():
    """src/texture.rs:1118-1139."""
    o = O.OracleScene(quad_scene(checker4()))
    mips = o.tex_levels(0, srgb=False)
    assert len(mips) == 3
    assert abs(mips[2][0, 0, 0] - 0.5) < 1e-6
    lod = o.tex_grad_lod(0, [[10.0, 0.0, 0.0, 10.0]])
    assert abs(o.tex_sample(0, [[0.1, 0.1]], lod)[0, 0] - 0.5) < 1e-5
    lod0 = o.tex_grad_lod(0, [[0.0, 0.0, 0.0, 0.0]])
    assert lod0[0] == 0.0
    assert abs(o.tex_sample(0, [[0.125, 0.125]], lod0)[0, 0] - 1.0) < 1e-6


def test_non_power_of_two_mips_terminate():
    """src/texture.rs:1141-1148."""
    o = O.OracleScene(quad_scene(np.full((3, 5, 4), 255, np.uint8)))
    mips = o.tex_levels(0, srgb=True)
    assert mips[-1].shape[:2] == (1, 1) and abs(mips[-1][0, 0, 0] - 1.0) < 1e-6
    assert [m.shape[:2] for m in mips] == [(3, 5), (1, 2), (1, 1)]


@pytest.mark.parametrize("mode,expect", [(A.MIN_NEAREST_MIP_NEAREST, "round"), (A.MIN_LINEAR_MIP_LINEAR, "lerp")])
def test_mip_modes_select_levels(mode, expect):
    """Level selection of sample_lod (src/texture.rs:417-447) on a texture
    whose levels are constant and distinct (8x8 of 1.0 except level 0 has
    a single bright texel pattern): NEAREST_MIP_NEAREST rounds the level,
    *_MIP_LINEAR blends two levels by the fraction."""
    px = np.zeros((8, 8, 4), np.uint8)
    px[::2, ::2] = 255  # level 1 = 0.25 everywhere, then 0.25 on
    o = O.OracleScene(quad_scene(px, dict(min_filter=mode)))
    mips = o.tex_levels(0)
    assert mips[1][0, 0, 0] == pytest.approx(0.25) and mips[3][0, 0, 0] == pytest.approx(0.25)
    v = o.tex_sample(0, [[0.3, 0.3]], 0.4)[0, 0]
    if expect == "round":
        # round(0.4) = 0: level 0 with the nearest texel filter.
        assert v == o.tex_sample(0, [[0.3, 0.3]], 0.0001)[0, 0]
    else:
        a = o.tex_sample(0, [[0.3, 0.3]], 0.0)[0, 0]  # level 0, bilinear (the magnification filter)
        b = 0.25
        assert v == pytest.approx(a + (b - a) * 0.4, abs=2e-6)


def test_cone_lod_closed_form():
    """TraceScene::lod (src/trace.rs:518-552) on the unit quad with uv = xy
    and a 256x256 texture, seen head-on: l = 0.5 log2(256^2) + log2(w) =
    8 + log2(w); with w = 1/256 the level is 0, with w = 1/64 it is 2 (and
    the cosine term: log2(1/cos) at 60 degrees adds 1)."""
    px = np.zeros((256, 256, 4), np.uint8)
    px[..., 0] = np.arange(256)[None, :]  # a ramp so levels differ
    px[..., 3] = 255
    sc = quad_scene(px, dict(mag=A.FILTER_LINEAR, min_filter=A.MIN_NEAREST_MIP_NEAREST))
    o = O.OracleScene(sc)
    bary = [[1 / 3, 1 / 3, 1 / 3]]
    mips = o.tex_levels(0, srgb=True)

    def base_at(width, d):
        return o.material_eval([0], bary, width=[width], dirs=[d])[0, 0]
    uv = np.array([[(0 + 1 + 1) / 3.0, (0 + 0 + 1) / 3.0]], np.float32)
    for w, dz, lvl in [(1 / 256, -1.0, 0), (1 / 64, -1.0, 2), (1 / 64, -0.5, 3)]:
        d = np.array([math.sqrt(1 - dz * dz), 0.0, dz], np.float32)
        got = base_at(w, d)
        # level 0 exactly: lod <= 0 selects the magnification filter.
        want = o.tex_sample(0, uv, float(lvl), srgb=True)[0, 0]
        assert got == pytest.approx(want, rel=1e-6), (w, dz, lvl)
    assert len(mips) == 9


def test_normal_map_tangent_frame():
    """TraceScene::perturb / tangent_frame (src/trace.rs:640-687): the
    normal-map texel (1, 0.5, 0.5) is tangent-space (1, 0, 0) x scale, so
    the mapped normal is the tangent; w < 0 flips the bitangent, so
    (0.5, 1, 0.5) maps to -bitangent; a flat texel keeps the normal."""
    def run(texel, w, scale=1.0, flip=0):
        px = np.zeros((1, 1, 4), np.float32)
        px[0, 0] = texel
        tang = np.tile(np.array([1.0, 0.0, 0.0, w], np.float32), (2, 3))
        sc = quad_scene(px, dict(mag=A.FILTER_NEAREST), tangents=tang, slots={"normal": 0}, normal_scale=scale)
        o = O.OracleScene(sc)
        return o.material_eval([0], [[0.2, 0.3, 0.5]], flip=[flip])[0, 5:8]
    np.testing.assert_allclose(run((1.0, 0.5, 0.5, 1.0), 1.0), [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(run((0.5, 1.0, 0.5, 1.0), 1.0), [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(run((0.5, 1.0, 0.5, 1.0), -1.0), [0, -1, 0], atol=1e-6)
    np.testing.assert_allclose(run((0.5, 0.5, 1.0, 1.0), 1.0), [0, 0, 1], atol=1e-6)
    # scale 0 leaves only the z part: the geometric normal.
    np.testing.assert_allclose(run((1.0, 0.5, 0.5, 1.0), 1.0, scale=0.0), [0, 0, 1], atol=1e-6)


def test_base_color_texture_and_vertex_colors_multiply():
    """material_oriented's base colour (src/trace.rs:724-735): factor x
    sRGB texel x interpolated COLOR_0."""
    px = np.zeros((1, 1, 4), np.uint8)
    px[0, 0] = (255, 128, 0, 255)
    cols = np.tile(np.array([0.5, 1.0, 1.0, 0.25], np.float32), (2, 3))
    mat = S.lambert((0.8, 0.6, 0.4))
    sc = quad_scene(px, dict(mag=A.FILTER_NEAREST), colors=cols, mat=mat)
    o = O.OracleScene(sc)
    r = o.material_eval([0], [[0.2, 0.3, 0.5]])[0]
    srgb128 = ((128 / 255 + 0.055) / 1.055) ** 2.4
    np.testing.assert_allclose(r[0:3], [0.8 * 1.0 * 0.5, 0.6 * srgb128 * 1.0, 0.0], rtol=1e-5, atol=1e-7)
    assert r[35] == pytest.approx(0.25)  # alpha: 1 x 1 x 0.25


def test_textured_emitter_power_uses_the_mean_texel():
    """EmissiveLights::build (src/pathtrace.rs:1484-1500): a textured
    emitter's power uses the emissive texture's 1x1 mip. Two equal quads,
    one with a half-black emissive texture: pmf 1/3 : 2/3."""
    def extra(s):
        e = s.add_material(S.lambert((0.0, 0.0, 0.0), emissive=(1.0, 1.0, 1.0)))
        e2 = s.add_material(S.lambert((0.0, 0.0, 0.0), emissive=(1.0, 1.0, 1.0)))
        p = np.array([[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float32)
        pos = np.stack([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])])
        uv = np.tile(np.array([0, 0, 1, 0, 1, 1], np.float32), (2, 1))
        s.add_tris(pos, e, uv0=uv)
        s.add_tris(pos + np.array([2, 0, 0] * 3, np.float32), e2, uv0=uv)
        half = np.zeros((2, 2, 4), np.float32)
        half[:, 0] = (1.0, 1.0, 1.0, 1.0)
        t = s.add_texture(half, mag=A.FILTER_NEAREST)
        s.set_material_textures(e, emissive=t)
    sc = quad_scene(extra=extra)
    o = O.OracleScene(sc)
    tris, pmf, cdf, lookup = o.emissive_table()
    assert len(tris) == 4
    np.testing.assert_allclose(pmf, [1 / 6, 1 / 6, 1 / 3, 1 / 3], rtol=1e-6)
    # Emission itself samples the texture at level 0: the left half lit.
    e = o.material_eval([2, 2], [[0.9, 0.05, 0.05], [0.05, 0.9, 0.05]])
    assert e[0, 39] == pytest.approx(1.0) and e[1, 39] == pytest.approx(0.0, abs=1e-7)


def masked_scene():
    """A Lambert floor (0.5) at z = -0.5, a MASK quad (1.0, cutoff 0.5) at
    z = 0 whose texture is transparent on its left half (u < 0.5), a
    4 cd point light high above and a camera looking down."""
    px = np.zeros((1, 2, 4), np.uint8)
    px[0, 0] = (255, 255, 255, 0)
    px[0, 1] = (255, 255, 255, 255)
    mat = S.lambert((1.0, 1.0, 1.0))
    mat.alpha_mode = A.OXR_ALPHA_MASK
    mat.alpha_cutoff = 0.5

    def extra(s):
        f = s.add_material(S.lambert((0.5, 0.5, 0.5)))
        p = np.array([[-1, -1, -0.5], [2, -1, -0.5], [2, 2, -0.5], [-1, 2, -0.5]], np.float32)
        s.add_tris(np.stack([np.concatenate([p[0], p[1], p[2]]), np.concatenate([p[0], p[2], p[3]])]), f)
        s.lights.append(A.light(A.OXR_LIGHT_POINT, intensity=400.0, position=(0.5, 0.5, 20.0),
                                direction=(0.0, 0.0, -1.0)))
    sc = quad_scene(px, dict(mag=A.FILTER_NEAREST), mat=mat, extra=extra)
    sc.camera = S.CameraSpec((0.5, 0.5, 3.0), (0.5, 0.5, 0.0), yfov_deg=12.0)
    sc.ambient = 0.0
    return sc


def test_alpha_mask_texture():
    """TraceScene::alpha / masked_out (src/trace.rs:492-515): base alpha x
    base-texture alpha at level 0 x COLOR_0 alpha, compared with the MASK
    cutoff by the any-hit filter (src/pathtrace.rs:1924-1934). Camera rays
    pass the transparent half and see the floor below it, lit through it
    (0.5 / pi x 400 / 20.5^2); the opaque half shows the quad (1 / pi x
    400 / 20^2)."""
    sc = masked_scene()
    o = O.OracleScene(sc)
    r = o.material_eval([0, 0], [[0.9, 0.05, 0.05], [0.05, 0.05, 0.9]])
    assert r[0, 42] == 0.0 and r[1, 42] == 1.0   # uv (0.1, 0.05) left, (0.9, ...) right
    W = 16
    img = o.render(W, W, 0, 4, O.options(W, W, max_bounces=1))
    m = img[..., 0] / img[..., 3]
    e_floor = 400.0 / 20.5 ** 2
    e_quad = 400.0 / 20.0 ** 2
    left, right = m[4:12, 1:6], m[4:12, 10:15]
    np.testing.assert_allclose(left, 0.5 / math.pi * e_floor, rtol=2e-2)
    np.testing.assert_allclose(right, 1.0 / math.pi * e_quad, rtol=2e-2)


def test_texture_transform_matches_the_extension():
    """KHR_texture_transform: uv' = T R S uv with R = [[c, s], [-s, c]]
    (the extension's matrix). Parity unpinned against the reference
    (oxideav-mesh3d is not vendored)."""
    px = np.zeros((1, 4, 4), np.float32)
    px[0, :, 0] = [0.1, 0.2, 0.3, 0.4]
    px[..., 3] = 1.0
    rot = math.pi / 2
    slots = {"base_color": dict(texture=0, offset=(0.5, 0.0), rotation=rot, scale=(2.0, 1.0))}
    sc = quad_scene(px, dict(mag=A.FILTER_NEAREST), slots=slots)
    o = O.OracleScene(sc)
    bary = np.array([[0.5, 0.3, 0.2]], np.float32)
    u = 0.5 * 0 + 0.3 * 1 + 0.2 * 1
    v = 0.5 * 0 + 0.3 * 0 + 0.2 * 1
    c, s = math.cos(rot), math.sin(rot)
    up = c * 2 * u + s * 1 * v + 0.5
    vp = -s * 2 * u + c * 1 * v
    want = o.tex_sample(0, [[up, vp]], 0.0, srgb=True)[0, 0]
    assert o.material_eval([0], bary)[0, 0] == pytest.approx(want, rel=1e-6)



def _cone_scene(projection):
    """The unit quad, white Lambert with a seeded random 256x256 float
    base-colour texture sampled trilinearly (so the base colour is a
    continuous function of the cone level), lit by one point light."""
    rng = np.random.default_rng(11)
    px = np.ones((256, 256, 4), np.float32)
    px[..., :3] = rng.random((256, 256, 3), dtype=np.float32)

    def light(s):
        s.lights.append(A.light(A.OXR_LIGHT_POINT, intensity=2.0, position=(0.5, 0.4, 1.0), direction=(0, 0, -1)))
    sc = quad_scene(px, dict(mag=A.FILTER_LINEAR, min_filter=A.MIN_LINEAR_MIP_LINEAR), extra=light)
    if projection == A.OXR_PROJ_ORTHOGRAPHIC:
        sc.camera = S.CameraSpec((0.45, 0.5, 2.0), (0.5, 0.52, 0.0), projection=projection, ymag=0.45)
    else:
        sc.camera = S.CameraSpec((0.45, 0.5, 1.6), (0.5, 0.52, 0.0), yfov_deg=30.0)
    sc.ambient = 0.0
    return sc


@pytest.mark.parametrize("projection", [A.OXR_PROJ_PERSPECTIVE, A.OXR_PROJ_ORTHOGRAPHIC])
def test_first_hit_cone_width_is_the_references(projection):
    """PathTracer::refine (src/pathtrace.rs:2537-2554) and the first hit of
    Integrator::sample (:2083-2092): the ray cone's width is
    spread = 2 half_h / height for an orthographic camera and spread * t for
    a perspective one. The only term in this scene is the point light's
    NEE on a white Lambert quad, base(w) / pi * I cos / d^2, so every
    covered sample's radiance divided by the base colour evaluated at that
    width (half_h from the camera definition here: tan(yfov / 2), or ymag)
    is the grey irradiance factor, the same in all channels and equal to
    its closed form."""
    W = H = 24
    sc = _cone_scene(projection)
    o = O.OracleScene(sc)
    opts = O.options(W, H, spp=1, max_bounces=1)
    rays = o.record_rays(W, H, 0, 1, 0, H, opts=opts)
    rays = rays[rays[:, 3] == 0]  # the primary rays, one per pixel
    assert len(rays) == W * H
    raw = o.trace_closest(O.RayBatch(rays[:, 0:3], rays[:, 4:7]).arr)
    prim = raw[:, 1]
    tuv = raw.view(np.float32)
    if projection == A.OXR_PROJ_ORTHOGRAPHIC:
        half_h = np.float32(0.45)
    else:
        half_h = np.float32(math.tan(math.radians(30.0) / 2.0))
    spread = np.float32(2.0) * half_h / np.float32(H)
    lp = np.array([0.5, 0.4, 1.0])
    checked = covered = 0
    for i in range(W * H):
        if prim[i] == 0xFFFFFFFF:
            continue
        covered += 1
        t, u, v = tuv[i, 0], tuv[i, 2], tuv[i, 3]
        width = spread if projection == A.OXR_PROJ_ORTHOGRAPHIC else np.float32(spread * t)
        bary = [[np.float32(1.0) - u - v, u, v]]
        base = o.material_eval([prim[i]], bary, width=[width], dirs=[rays[i, 4:7]])[0, 0:3]
        got = o.sample_one(W, H, i % W, i // W, 0, opts=opts)
        p = rays[i, 0:3].astype(np.float64) + float(t) * rays[i, 4:7].astype(np.float64)
        dl = lp - p
        d2 = float(dl @ dl)
        k = 2.0 * (dl[2] / math.sqrt(d2)) / (math.pi * d2)
        np.testing.assert_allclose(got / base, [k, k, k], rtol=2e-5, err_msg=f"pixel {i}")
        # A cone of zero width (mip level 0) gives a different base colour.
        lvl0 = o.material_eval([prim[i]], bary, width=[0.0], dirs=[rays[i, 4:7]])[0, 0:3]
        checked += int(not np.allclose(lvl0, base, rtol=1e-3))
    assert covered > 0.9 * W * H and checked > 0.9 * covered
