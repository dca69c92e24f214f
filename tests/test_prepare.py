"""The prepare layer (src/prepare.rs:1098-1575) on the CPU: the
reference's own prepare tests (src/prepare.rs:1599-1776), the glTF rules
that fix strips, fans and T * R * S, closed forms for tangents and the
normal matrix, the material and texture-binding rules, prepared lights
checked through the oracle's integrator (irradiance of a Lambert plane
under the prepared spot and fallback lights), and the animated pipeline:
the reference's skinned, morphed beam (src/testscenes.rs:576-680) in
closed form at t = 0, 0.5 and 1 with the coverage its scanline test
asserts (tests/scanline_pbr.rs:252-270), and Appendix C's samplers,
morph weight precedence, skinned normals and variants against closed
forms."""
import math

import numpy as np
import pytest

from oxr import _abi as A
from oxr import prepare as P
from oxr import scenes


def tri_scene(**node_kw):
    """src/prepare.rs:1585-1593: one triangle in z = 0 under a node."""
    prim = P.Primitive(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))
    s = P.Scene3D(meshes=[P.Mesh([prim])], nodes=[P.Node(mesh=0, **node_kw)], roots=[0])
    return s


def face_z(t):
    return np.cross(t[1] - t[0], t[2] - t[0])[2]


def test_flat_normals_and_world_space_positions():
    p = P.build(tri_scene(translation=(0.0, 0.0, 2.0)))
    assert len(p.items) == 1
    it = p.items[0]
    assert list(it.positions[1]) == [1.0, 0.0, 2.0]
    assert not it.has_vertex_normals
    assert list(it.normals[0]) == [0.0, 0.0, 1.0]
    assert it.material == p.default_material()
    assert len(p.lights) == 1 and not p.scene_lights
    assert list(p.bounds[1]) == [1.0, 1.0, 2.0]
    fb = p.materials[p.default_material()]
    assert list(fb.base_color) == pytest.approx([0.7, 0.7, 0.75, 1.0])
    assert fb.metallic == 0.0 and fb.roughness == 1.0


def test_mirrored_world_keeps_ccw_front_facing():
    p = P.build(tri_scene(scale=(-1.0, 1.0, 1.0)))
    assert face_z(p.items[0].positions) > 0.0
    # The unflipped winding would face -Z.
    q = P.build(tri_scene())
    assert face_z(q.items[0].positions) > 0.0
    assert list(p.items[0].positions[1]) == [0.0, 1.0, 0.0]  # (a, c, b)


def test_strips_lines_and_garbage_indices():
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    strip = P.Primitive(positions=quad, topology="triangle_strip")
    loop = P.Primitive(positions=quad, topology="line_loop")
    bad = P.Primitive(positions=quad[:3], indices=np.array([0, 1, 2, 0, 1, 99], np.uint32))
    s = P.Scene3D(meshes=[P.Mesh([strip, loop, bad])], nodes=[P.Node(mesh=0)], roots=[0])
    p = P.build(s)
    assert len(p.items) == 3
    assert p.items[0].primitive_count() == 2
    assert p.items[1].topology == "lines" and p.items[1].primitive_count() == 4
    assert p.items[2].primitive_count() == 1
    for t in p.items[0].positions.reshape(-1, 3, 3):
        assert face_z(t) > 0.0
    # Line loop closes back to the first vertex.
    assert list(p.items[1].positions[-1]) == list(quad[0])
    # Lines trace nothing: the tracer's soup holds the triangles only.
    assert P.to_scene(p).num_tris == 3


def test_strip_and_fan_follow_gltf_triangulation():
    # glTF 2.0 3.7.2.1: strip i -> (i, i+1+i%2, i+2-i%2); fan i -> (i+1, i+2, 0).
    ring = np.array([[0, 0, 0]] + [[math.cos(a), math.sin(a), 0.0] for a in np.linspace(0, 1.5 * math.pi, 5)],
                    np.float32)
    fan = P.Primitive(positions=ring, topology="triangle_fan")
    strip = P.Primitive(positions=ring[[1, 2, 0, 3, 4]], topology="triangle_strip", indices=None)
    tris = P.triangle_indices(strip, list(range(5)))
    assert tris == [(0, 1, 2), (1, 3, 2), (2, 3, 4)]
    assert P.triangle_indices(fan, list(range(6))) == [(1, 2, 0), (2, 3, 0), (3, 4, 0), (4, 5, 0)]
    p = P.build(P.Scene3D(meshes=[P.Mesh([fan])], nodes=[P.Node(mesh=0)], roots=[0]))
    assert p.items[0].primitive_count() == 4
    for t in p.items[0].positions.reshape(-1, 3, 3):
        assert face_z(t) > 0.0
    # Fewer than three vertices make no triangle and no item.
    short = P.Primitive(positions=ring[:2], topology="triangle_strip")
    assert P.build(P.Scene3D(meshes=[P.Mesh([short])], nodes=[P.Node(mesh=0)], roots=[0])).items == []


def test_non_finite_positions_drop_their_triangles():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [np.nan, 0, 0]], np.float32)
    prim = P.Primitive(positions=pos, indices=np.array([0, 1, 2, 0, 3, 2], np.uint16))
    p = P.build(P.Scene3D(meshes=[P.Mesh([prim])], nodes=[P.Node(mesh=0)], roots=[0]))
    assert p.items[0].primitive_count() == 1


def test_scene_lights_resolve_with_world_transforms(oracle):
    s = tri_scene()
    s.lights.append(P.Light("spot", color=(1.0, 0.5, 0.25), intensity=10.0, range=5.0, inner_cone_angle=0.2,
                            outer_cone_angle=0.4))
    h = math.sqrt(0.5)
    s.nodes.append(P.Node(light=0, translation=(0.0, 3.0, 0.0), rotation=(-h, 0.0, 0.0, h)))
    s.roots.append(1)
    p = P.build(s)
    assert p.scene_lights
    l = p.lights[0]
    assert l.kind == A.OXR_LIGHT_SPOT
    assert abs(l.direction[1] + 1.0) < 1e-5, list(l.direction)
    assert list(l.position) == pytest.approx([0.0, 3.0, 0.0])
    assert l.range == 5.0
    assert l.inner_cone_angle == pytest.approx(0.2) and l.outer_cone_angle == pytest.approx(0.4)
    # Irradiance through the integrator: a white Lambert plane reflects
    # E / pi. Straight below at y = 1: 10 / 2^2 * (1 - (2/5)^4).
    window = 1.0 - (2.0 / 5.0) ** 4
    below = lambert_radiance(oracle, l, (0.0, 1.0, 0.0))
    assert below[0] == pytest.approx(10.0 / 4.0 * window / math.pi, rel=1e-4)
    assert below[1] == pytest.approx(0.5 * below[0], rel=1e-5)
    assert np.all(lambert_radiance(oracle, l, (3.0, 2.0, 0.0)) == 0.0)   # outside the outer cone
    assert np.all(lambert_radiance(oracle, l, (0.0, -3.0, 0.0)) == 0.0)  # beyond range
    p2 = P.build(s, use_scene_lights=False)
    assert not p2.scene_lights
    assert p2.lights[0].kind == A.OXR_LIGHT_DIRECTIONAL


def lambert_radiance(oracle, light, at):
    """Mean direct radiance of a small white Lambert patch at `at` facing +Y
    lit by `light` alone (a one-bounce render of the patch)."""
    x, y, z = at
    s = scenes.Scene("patch")
    m = s.add_material(scenes.lambert((1.0, 1.0, 1.0)))
    e = 1.0e-3
    t, n = scenes.quad([[x - e, y, z + e], [x + e, y, z + e], [x + e, y, z - e], [x - e, y, z - e]])
    s.add_tris(t, m, normals=n)
    s.lights.append(light)
    s.camera = scenes.CameraSpec((x, y + 0.5, z + 1e-4), (x, y, z), yfov_deg=0.01)
    s.ambient = 0.0
    os_ = oracle.OracleScene(s)
    acc = os_.render(2, 2, 0, 1, oracle.options(2, 2, spp=1, max_bounces=1))
    assert np.all(acc[..., 3] == 1.0)
    return acc[..., :3].reshape(-1, 3).mean(0)


def test_fallback_light_matches_legacy_lambert_scale(oracle):
    l = P.light_from_spec(P.LightSpec(azimuth_deg=0.0, elevation_deg=90.0, intensity=1.0))
    assert l.kind == A.OXR_LIGHT_DIRECTIONAL
    assert -l.direction[1] == pytest.approx(1.0, abs=1e-6)
    assert l.intensity == pytest.approx(math.pi)
    # E / pi = 1: a white Lambert plane facing the light reflects 1.
    assert lambert_radiance(oracle, l, (0.0, 0.0, 0.0)) == pytest.approx([1.0, 1.0, 1.0], rel=1e-5)
    d = P.light_from_spec(P.LightSpec())  # az 45, el 45
    assert list(d.direction) == pytest.approx([-0.5, -math.sqrt(0.5), -0.5], abs=1e-6)
    assert P.light_from_spec(P.LightSpec(intensity=-3.0)).intensity == 0.0
    assert P.build(tri_scene(), fallback_light=None).lights == []


def test_light_sanitising():
    s = tri_scene()
    s.lights += [P.Light("spot", intensity=float("nan"), color=(-1.0, float("inf"), 2.0),
                         outer_cone_angle=3.0, inner_cone_angle=5.0, range=-1.0),
                 P.Light("spot", outer_cone_angle=float("nan"), inner_cone_angle=float("nan")),
                 P.Light("point", range=float("inf")), P.Light("directional", range=4.0)]
    for i in range(4):
        s.nodes.append(P.Node(light=i, scale=(0.0, 0.0, 0.0)))  # degenerate: default direction
        s.roots.append(i + 1)
    L = P.build(s).lights
    assert L[0].intensity == 0.0 and list(L[0].color) == [0.0, 0.0, 2.0]
    assert L[0].outer_cone_angle == pytest.approx(math.pi / 2) and L[0].inner_cone_angle == pytest.approx(math.pi / 2)
    assert L[0].range == 0.0
    assert L[1].outer_cone_angle == pytest.approx(math.pi / 4) and L[1].inner_cone_angle == 0.0
    assert L[2].range == 0.0 and L[2].outer_cone_angle == 0.0
    assert L[3].range == 0.0
    assert all(list(l.direction) == [0.0, 0.0, -1.0] for l in L)


def test_camera_instances_collected():
    s = tri_scene()
    s.cameras.append(P.Camera(yfov=0.8, znear=0.1))
    s.nodes.append(P.Node(camera=0, translation=(0.0, 1.0, 5.0)))
    s.roots.append(1)
    p = P.build(s)
    assert len(p.cameras) == 1
    node, ci, cam, world = p.cameras[0]
    assert node == 1 and ci == 0 and cam.yfov == 0.8
    assert list(world[:3, 3]) == [0.0, 1.0, 5.0]


def test_normal_matrix_of_scale_is_inverse():
    m = np.eye(4, dtype=np.float32)
    m[0, 0] = 2.0
    n = P.normal_matrix(m)
    assert abs(n[0, 0] - 0.5) < 1e-6 and abs(n[1, 1] - 1.0) < 1e-6
    rng = np.random.default_rng(1)
    for _ in range(20):
        m = np.eye(4, dtype=np.float32)
        m[:3, :3] = rng.normal(size=(3, 3))
        np.testing.assert_allclose(P.normal_matrix(m), np.linalg.inv(m[:3, :3].astype(np.float64)).T, rtol=2e-4,
                                   atol=2e-4)
    assert np.array_equal(P.normal_matrix(np.zeros((4, 4), np.float32)), np.eye(3))


def test_node_hierarchy_composes_trs_in_preorder():
    # Parent: translate (1, 0, 0); child: rotate 90 deg about +Z, scale 2.
    h = math.sqrt(0.5)
    prim = P.Primitive(positions=np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32))
    s = P.Scene3D(meshes=[P.Mesh([prim])], roots=[0, 0, 3],
                  nodes=[P.Node(children=[1, 2], translation=(1.0, 0.0, 0.0)),
                         P.Node(mesh=0, rotation=(0.0, 0.0, h, h), scale=(2.0, 2.0, 2.0)),
                         P.Node(mesh=0, children=[1]),   # shared child: visited once
                         P.Node(mesh=7)])                 # bad mesh index: skipped
    assert P.preorder_nodes(s) == [0, 1, 2, 3]
    p = P.build(s)
    assert [it.node for it in p.items] == [1, 2]
    np.testing.assert_allclose(p.items[0].positions, [[1, 2, 0], [-1, 0, 0], [1, 0, 2]], atol=1e-6)
    # A matrix node takes the matrix as given.
    m = np.eye(4)
    m[:3, 3] = (0.0, 5.0, 0.0)
    s2 = P.Scene3D(meshes=[P.Mesh([prim])], roots=[0], nodes=[P.Node(mesh=0, matrix=m, translation=(9, 9, 9))])
    np.testing.assert_allclose(P.build(s2).items[0].positions[0], [1, 5, 0])


def test_vertex_normals_follow_the_inverse_transpose():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[1, 1, 0]] * 3, np.float32)
    prim = P.Primitive(positions=pos, normals=nrm)
    p = P.build(P.Scene3D(meshes=[P.Mesh([prim])], roots=[0], nodes=[P.Node(mesh=0, scale=(2.0, 1.0, 1.0))]))
    it = p.items[0]
    assert it.has_vertex_normals
    want = np.array([0.5, 1.0, 0.0]) / math.hypot(0.5, 1.0)
    np.testing.assert_allclose(it.normals, [want] * 3, atol=1e-6)
    # A wrong-length normal stream falls back to flat normals.
    bad = P.Primitive(positions=pos, normals=nrm[:2])
    q = P.build(P.Scene3D(meshes=[P.Mesh([bad])], roots=[0], nodes=[P.Node(mesh=0)]))
    assert not q.items[0].has_vertex_normals


def test_face_tangents_closed_form():
    # uv = (x, y) on a +Z triangle: tangent +X, bitangent +Y, w = +1;
    # mirroring u flips the handedness.
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 1]] * 3, np.float32)
    t = P.face_tangents(pos, nrm, np.array([[0, 0], [1, 0], [0, 1]], np.float32))
    np.testing.assert_allclose(t, [[1, 0, 0, 1]] * 3, atol=1e-6)
    t = P.face_tangents(pos, nrm, np.array([[0, 0], [-1, 0], [0, 1]], np.float32))
    np.testing.assert_allclose(t, [[-1, 0, 0, -1]] * 3, atol=1e-6)
    # Degenerate UVs: default frame, Gram-Schmidt against the normal.
    t = P.face_tangents(pos, np.array([[1, 0, 0]] * 3, np.float32), np.zeros((3, 2), np.float32))
    assert np.all(np.abs(np.einsum("ij,j->i", t[:, :3], [1, 0, 0])) < 1e-6)
    np.testing.assert_allclose(np.linalg.norm(t[:, :3], axis=1), 1.0, atol=1e-6)


def test_tangents_generated_only_for_normal_mapped_materials():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    uv = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    tex = P.Texture(np.full((2, 2, 4), 128, np.uint8))
    mats = [P.Material(textures={"normal": P.TextureRef(0, tex_coord=1)}), P.Material(),
            P.Material(textures={"normal": P.TextureRef(5)})]  # missing texture: not bound
    prims = [P.Primitive(positions=pos, uvs=[np.zeros((3, 2), np.float32), uv], material=0),
             P.Primitive(positions=pos, uvs=[uv], material=1),
             P.Primitive(positions=pos, uvs=[uv], material=2),
             P.Primitive(positions=pos, uvs=[uv], material=9)]
    s = P.Scene3D(meshes=[P.Mesh(prims)], roots=[0], nodes=[P.Node(mesh=0)], materials=mats, textures=[tex])
    p = P.build(s)
    np.testing.assert_allclose(p.items[0].tangents, [[1, 0, 0, 1]] * 3, atol=1e-6)  # from TEXCOORD_1
    assert [len(it.tangents) for it in p.items[1:]] == [0, 0, 0]
    assert [it.material for it in p.items] == [0, 1, 2, 3]
    assert len(P.build(s, generate_tangents=False).items[0].tangents) == 0
    # Given tangents are normalised, with w reduced to its sign.
    given = P.Primitive(positions=pos, uvs=[uv], material=0,
                        tangents=np.array([[0, 3, 0, -0.2]] * 3, np.float32))
    s.meshes[0].primitives = [given]
    np.testing.assert_allclose(P.build(s).items[0].tangents, [[0, 1, 0, -1]] * 3)


def test_material_rules():
    m = P.Material(base_color=(float("nan"), 0.5, 2.0, 1.5), metallic=float("nan"), roughness=-1.0,
                   emissive_factor=(1.0, -2.0, float("nan")), emissive_strength=4.0, ior=0.5,
                   double_sided=True, alpha_mode=A.OXR_ALPHA_MASK, alpha_cutoff=0.3,
                   params=dict(transmission=0.25))
    q = P.prepared_material(m)
    assert list(q.base_color) == [1.0, 0.5, 2.0, 1.0]
    assert q.metallic == 1.0 and q.roughness == 0.0
    assert list(q.emissive) == [4.0, 0.0, 0.0]
    assert q.ior == 1.0
    assert q.flags & A.OXR_MAT_DOUBLE_SIDED
    assert q.alpha_mode == A.OXR_ALPHA_MASK and q.alpha_cutoff == pytest.approx(0.3)
    assert q.transmission == 0.25
    assert P.prepared_material(P.Material()).ior == 1.5
    assert P.prepared_material(P.Material(ior=float("nan"), emissive_strength=-1.0,
                                          emissive_factor=(1.0, 1.0, 1.0))).ior == 1.5
    assert list(P.prepared_material(P.Material(emissive_strength=-1.0, emissive_factor=(1, 1, 1))).emissive) \
        == [0.0, 0.0, 0.0]


def test_texture_binding_transform_filter():
    tex = P.Texture(np.zeros((1, 1, 4), np.uint8))
    refs = {"base_color": P.TextureRef(0, offset=(0.0, 0.0), rotation=0.0, scale=(1.0, 1.0)),
            "emissive": P.TextureRef(0, tex_coord=0, transform_tex_coord=1, offset=(0.5, 0.0)),
            "metallic_roughness": P.TextureRef(0, scale=(float("inf"), 1.0)),
            "occlusion": P.TextureRef(0)}
    s = P.Scene3D(meshes=[P.Mesh([P.Primitive(positions=np.eye(3, dtype=np.float32), material=0)])],
                  roots=[0], nodes=[P.Node(mesh=0)], materials=[P.Material(textures=refs)], textures=[tex])
    sc = P.to_scene(P.build(s))
    mt = sc.material_tex[0]
    assert mt.base_color.texture == 0 and mt.base_color.has_transform == 0
    assert mt.emissive.uv_set == 1 and mt.emissive.has_transform == 1 and mt.emissive.offset[0] == 0.5
    assert mt.metallic_roughness.has_transform == 0
    assert mt.normal.texture == -1


def test_to_scene_packs_items_in_order():
    quad = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], np.float32)
    cols = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 1, 1]], np.float32)
    uv = quad[:, :2].copy()
    a = P.Primitive(positions=quad, topology="triangle_strip", colors=[cols], uvs=[uv, uv * 2])
    b = P.Primitive(positions=quad[:3], normals=np.array([[0, 0, 2]] * 3, np.float32), material=0)
    s = P.Scene3D(meshes=[P.Mesh([a, b])], roots=[0, 1],
                  nodes=[P.Node(mesh=0), P.Node(mesh=0, translation=(0.0, 0.0, 1.0))],
                  materials=[P.Material(base_color=(0.2, 0.4, 0.6, 1.0))])
    p = P.build(s)
    sc = P.to_scene(p)
    assert sc.num_tris == 6
    assert sc.num_materials == 2
    np.testing.assert_array_equal(sc.material_ids, [1, 1, 0, 1, 1, 0])
    pos, nrm, flags, _ = sc.pack()
    np.testing.assert_array_equal(pos[1], [1, 0, 0, 1, 1, 0, 0, 1, 0])  # strip triangle 1: (1, 3, 2)
    np.testing.assert_array_equal(pos[3, 2::3], [1, 1, 1])
    assert flags[0] == A.TRI_HAS_UV0 | A.TRI_HAS_UV1 | A.TRI_HAS_COLORS
    assert flags[2] == A.OXR_TRI_HAS_NORMALS
    np.testing.assert_array_equal(nrm[2], [0, 0, 1] * 3)
    at = sc.attributes()
    np.testing.assert_array_equal(at["colors"][1], np.concatenate([cols[1], cols[3], cols[2]]))
    np.testing.assert_array_equal(at["uv1"][0], (uv[[0, 1, 2]] * 2).ravel())


def test_showcase_graph_prepares_outward_boxes_and_generated_tangents():
    g = P.showcase_graph()
    p = P.build(g)
    assert p.scene_lights and len(p.lights) == 2
    assert [it.topology for it in p.items] == ["triangles", "lines", "triangles", "triangles", "triangles",
                                               "triangles", "triangles"]
    assert [it.primitive_count() for it in p.items] == [2, 4, 12, 12, 2, 8, 1]
    for k in (2, 3):  # both box instances, one under a mirroring node
        it = p.items[k]
        centre = (it.bounds_min + it.bounds_max) / 2
        t = it.positions.reshape(-1, 3, 3)
        n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
        assert np.all(np.einsum("ij,ij->i", n, t.mean(1) - centre) > 0.0), k
        assert len(it.colors) == 36
    assert p.materials[0].base_color[0] == pytest.approx(0.8)
    panel = p.items[4]
    assert len(panel.tangents) == 6
    # u runs along +x and v down the panel (-y), so B = -y = w * (N x T)
    # with N x T = +y: w = -1.
    np.testing.assert_allclose(panel.tangents, [[1, 0, 0, -1]] * 6, atol=1e-6)
    lamp = p.items[5]
    assert np.all(lamp.normals[:, 1] < -0.99)  # the fan faces down after the node's half turn
    assert p.items[6].material == p.default_material()
    sc = P.to_scene(p)
    assert sc.num_tris == 37
    assert list(sc.material(3).emissive) == pytest.approx([6.0, 4.8, 3.6])


def _vec(c, name):
    return np.array(list(getattr(c, name)), np.float64)


def test_scene_camera_from_its_node():
    """Camera::from_scene_camera (src/camera.rs:208-332): eye and axes from
    the node's world matrix with its scale ignored, the viewport aspect,
    sanitised yfov and ymag."""
    world = P.local_matrix(P.Node(translation=(0.0, 1.0, 5.0), scale=(2.0, 2.0, 2.0))).astype(np.float32)
    c = P.camera_from_instance(P.Camera(yfov=0.8, aspect=1.0), world, 200, 100)
    np.testing.assert_allclose(_vec(c, "eye"), [0, 1, 5])
    np.testing.assert_allclose(_vec(c, "forward"), [0, 0, -1])
    np.testing.assert_allclose(_vec(c, "side"), [1, 0, 0])
    np.testing.assert_allclose(_vec(c, "up"), [0, 1, 0])
    assert c.half_h == pytest.approx(math.tan(0.4), rel=1e-6)
    assert c.half_w == pytest.approx(2.0 * math.tan(0.4), rel=1e-6)
    assert c.projection == A.OXR_PROJ_PERSPECTIVE
    for bad in (0.0, float("nan"), 3.2):
        assert P.camera_from_instance(P.Camera(yfov=bad), world, 1, 1).half_h == pytest.approx(math.tan(math.pi / 6))
    o = P.camera_from_instance(P.Camera(perspective=False, ymag=-2.0), world, 100, 100)
    assert o.projection == A.OXR_PROJ_ORTHOGRAPHIC and o.half_h == 2.0 and o.half_w == 2.0
    assert P.camera_from_instance(P.Camera(perspective=False, ymag=0.0), world, 100, 50).half_w == 2.0
    # A quarter turn about +Y: the camera's -Z looks down -X.
    h = math.sqrt(0.5)
    r = P.local_matrix(P.Node(rotation=(0.0, h, 0.0, h))).astype(np.float32)
    c = P.camera_from_instance(P.Camera(), r, 64, 64)
    np.testing.assert_allclose(_vec(c, "forward"), [-1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(_vec(c, "side"), [0, 0, -1], atol=1e-6)
    # Looking straight down: the node's +Y up hint keeps the frame valid.
    d = P.local_matrix(P.Node(rotation=(-h, 0.0, 0.0, h))).astype(np.float32)
    c = P.camera_from_instance(P.Camera(), d, 64, 64)
    np.testing.assert_allclose(_vec(c, "forward"), [0, -1, 0], atol=1e-6)
    np.testing.assert_allclose(_vec(c, "up"), [0, 0, -1], atol=1e-6)


def test_auto_frame_and_orbit():
    """Camera::build (src/camera.rs:357-475): distance radius / tan(fov/2)
    with radius 0.6 * extent, looking down -Z; an orbit multiplies that
    distance; orthographic frames 1.2x the half extent on the smaller axis."""
    b = ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    c = P.frame_bounds(100, 100, b)
    dist = 1.2 / math.tan(math.radians(30.0))
    np.testing.assert_allclose(_vec(c, "eye"), [0, 0, dist], rtol=1e-6)
    np.testing.assert_allclose(_vec(c, "forward"), [0, 0, -1])
    assert c.half_h == pytest.approx(math.tan(math.radians(30.0)), rel=1e-6)
    o = P.frame_bounds(100, 100, b, orbit=(30.0, 90.0, 2.0), target_offset=(0.0, 1.0, 0.0))
    want = np.array([0, 1, 0]) + 2 * dist * np.array([math.cos(math.radians(30)), 0.5, 0.0])
    np.testing.assert_allclose(_vec(o, "eye"), want, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(_vec(o, "forward"), -(want - [0, 1, 0]) / np.linalg.norm(want - [0, 1, 0]), atol=1e-6)
    w = P.frame_bounds(200, 100, b, projection=A.OXR_PROJ_ORTHOGRAPHIC)
    assert (w.half_w, w.half_h) == pytest.approx((2.4, 1.2))
    t = P.frame_bounds(100, 200, b, projection=A.OXR_PROJ_ORTHOGRAPHIC, orbit=(0.0, 0.0, 0.5))
    assert (t.half_w, t.half_h) == pytest.approx((0.6, 1.2))
    # Camera::resolve: the named scene camera, else the bounds (a unit box
    # when there is no geometry).
    g = P.showcase_graph()
    prep = P.build(g)
    assert len(prep.cameras) == 1
    pc = P.PreparedCamera(prep, scene_camera=0).resolve(64, 64)
    np.testing.assert_allclose(_vec(pc, "eye"), [0, 1.7, 3.4], atol=1e-6)
    assert _vec(pc, "forward")[1] < -0.29
    fb = P.PreparedCamera(prep, scene_camera=5).resolve(64, 64)  # no such camera: auto-frame
    centre = (prep.bounds[0] + prep.bounds[1]) / 2
    np.testing.assert_allclose(_vec(fb, "eye")[:2], centre[:2], atol=1e-6)
    empty = P.PreparedCamera(P.build(P.Scene3D())).resolve(10, 10)
    np.testing.assert_allclose(_vec(empty, "eye"), [0, 0, 0.6 / math.tan(math.radians(30))], rtol=1e-6)


# ---------------------------------------------------------------------------
# Animation, morph targets, skinning and variants.
# ---------------------------------------------------------------------------
def test_animation_time_moves_geometry():
    """src/prepare.rs:1735-1756."""
    s = tri_scene()
    s.animations = [P.Animation(channels=[P.Channel(0, "translation", P.AnimationSampler(
        np.array([0.0, 1.0]), np.array([[0.0, 0.0, 0.0], [10.0, 0.0, 0.0]])))])]
    rest = P.build(s)
    assert list(rest.items[0].positions[0]) == [0.0, 0.0, 0.0]
    half = P.build(s, time=0.5)
    assert abs(half.items[0].positions[0][0] - 5.0) < 1e-5
    # A non-finite time, or an animation index out of range, is the rest pose.
    for kw in (dict(time=float("nan")), dict(time=0.5, animation=3)):
        assert list(P.build(s, **kw).items[0].positions[0]) == [0.0, 0.0, 0.0]


def test_sampler_interpolation_closed_forms():
    S = P.AnimationSampler
    keys = np.array([1.0, 2.0, 4.0])
    vals = np.array([[0.0], [10.0], [30.0]])
    lin = S(keys, vals)
    # Clamped outside the keys; linear inside each interval.
    for t, want in ((0.0, 0.0), (1.0, 0.0), (1.25, 2.5), (3.0, 20.0), (4.0, 30.0), (9.0, 30.0)):
        assert P.sample_channel(lin, t)[0] == pytest.approx(want, abs=1e-12)
    step = S(keys, vals, "STEP")
    assert [P.sample_channel(step, t)[0] for t in (1.5, 2.0, 3.99, 4.5)] == [0.0, 10.0, 10.0, 30.0]
    # Cubic Hermite reproduces a cubic from its values and derivatives:
    # f(t) = t^3 on [0, 2], f'(0) = 0, f'(2) = 12 (tangents are per unit time).
    cub = S(np.array([0.0, 2.0]), np.array([[5.0], [0.0], [0.0], [12.0], [8.0], [7.0]]), "CUBICSPLINE")
    for t in (0.0, 0.5, 1.0, 1.7, 2.0):
        assert P.sample_channel(cub, t)[0] == pytest.approx(t ** 3, abs=1e-12)
    # Slerp from identity to 90 degrees about Z: 45 degrees at u = 0.5 and
    # 22.5 degrees at u = 0.25 (not the normalised lerp, which gives 26.6).
    h = math.sqrt(0.5)
    rot = S(np.array([0.0, 1.0]), np.array([[0.0, 0.0, 0.0, 1.0], [0.0, 0.0, h, h]]))
    for u, ang in ((0.5, 45.0), (0.25, 22.5)):
        q = P.sample_channel(rot, u, rotation=True)
        a = math.radians(ang) / 2
        assert q == pytest.approx([0.0, 0.0, math.sin(a), math.cos(a)], abs=1e-12)
    # The shorter arc: -q is the same rotation, and the result does not change.
    rot_neg = S(np.array([0.0, 1.0]), np.array([[0.0, 0.0, 0.0, 1.0], [0.0, 0.0, -h, -h]]))
    assert P.sample_channel(rot_neg, 0.5, rotation=True) == pytest.approx(P.sample_channel(rot, 0.5, True),
                                                                          abs=1e-12)
    # Malformed outputs (not a whole row count per key) sample to nothing.
    assert P.sample_channel(S(np.array([0.0, 1.0]), np.zeros(3)), 0.5) is None


def _beam_vertex(prep, x, y):
    """Prepared position of the beam's rest vertex (x, y)."""
    scene = P.skinned_morph_beam()
    rest = scene.meshes[0].primitives[0].positions
    idx = np.asarray(scene.meshes[0].primitives[0].indices)
    it = prep.items[0]
    k = int(np.argmin(np.abs(rest[:, 0] - x) + np.abs(rest[:, 1] - y)))
    return it.positions[int(np.nonzero(idx == k)[0][0])]


def _covered(prep, x, y):
    """Whether (x, y) in z = 0 lies inside one of the prepared triangles."""
    t = prep.items[0].positions.reshape(-1, 3, 3)[:, :, :2].astype(np.float64)
    p = np.array([x, y])

    def cross(a, b, c):
        return (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    pp = np.broadcast_to(p, (len(t), 2))
    s0, s1, s2 = cross(t[:, 0], t[:, 1], pp), cross(t[:, 1], t[:, 2], pp), cross(t[:, 2], t[:, 0], pp)
    return bool(np.any(((s0 >= 0) & (s1 >= 0) & (s2 >= 0)) | ((s0 <= 0) & (s1 <= 0) & (s2 <= 0))))


def test_skinned_morph_beam_closed_form():
    """The reference's beam (src/testscenes.rs:576-680). At time t the elbow
    has turned 90t degrees about +Z and the morph lifts by 0.3t, so a
    vertex of rest position (x, y) with elbow weight w lands at
    (1 - w) p + w (E R (p - E)) with p = (x, y + 0.3t) and E = (1, 0)."""
    scene = P.skinned_morph_beam()
    for t in (0.0, 0.5, 1.0):
        prep = P.build(scene, time=t)
        assert len(prep.items) == 1 and prep.items[0].skinned
        assert np.array_equal(prep.items[0].world, np.eye(4, dtype=np.float32))
        a = math.radians(90.0 * t)
        for x, y in ((2.0, 0.1), (2.0, -0.1), (1.0, 0.1), (0.5, -0.1), (0.3, 0.1), (1.6, -0.1)):
            w = min(max(x - 0.5, 0.0), 1.0)
            px, py = x, y + 0.3 * t
            bx, by = math.cos(a) * (px - 1.0) - math.sin(a) * py + 1.0, math.sin(a) * (px - 1.0) + math.cos(a) * py
            want = ((1 - w) * px + w * bx, (1 - w) * py + w * by, 0.0)
            assert _beam_vertex(prep, x, y) == pytest.approx(want, abs=2e-6), (t, x, y)
        # Normals stay +Z: a rotation about Z, and scale-free blends.
        assert np.allclose(prep.items[0].normals, [0.0, 0.0, 1.0], atol=1e-6)
    # tests/scanline_pbr.rs:252-270: the rest tip is covered at t = 0 and
    # left at t = 1, where the bent and morphed tip is covered instead.
    p0, p1 = P.build(scene, time=0.0), P.build(scene, time=1.0)
    assert _covered(p0, 1.9, 0.0) and not _covered(p0, 0.72, 0.95)
    assert not _covered(p1, 1.9, 0.0) and _covered(p1, 0.72, 0.95)
    # Without a time the rest pose is the t = 0 pose here.
    assert np.array_equal(P.build(scene).items[0].positions, p0.items[0].positions)
    # A skinned mesh ignores its node's transform (glTF 3.7.3.2).
    scene.nodes[2].translation = (5.0, 5.0, 5.0)
    assert np.array_equal(P.build(scene, time=1.0).items[0].positions, p1.items[0].positions)


def test_morph_weight_precedence_and_linearity():
    prim = P.Primitive(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                       targets=[P.MorphTarget(positions=np.array([[0, 0, 1]] * 3, np.float32)),
                                P.MorphTarget(positions=np.array([[1, 0, 0], [0, 0, 0], [0, 2, 0]], np.float32))])
    s = P.Scene3D(meshes=[P.Mesh([prim], weights=[0.5, 0.25])], nodes=[P.Node(mesh=0)], roots=[0])
    base = prim.positions.astype(np.float64)
    d0, d1 = prim.targets[0].positions, prim.targets[1].positions

    def first(prep):
        return prep.items[0].positions
    assert np.allclose(first(P.build(s)), base + 0.5 * d0 + 0.25 * d1)           # the mesh's weights
    s.nodes[0].weights = [1.0, -1.0]
    assert np.allclose(first(P.build(s)), base + d0 - d1)                        # the node's override them
    s.animations = [P.Animation(channels=[P.Channel(0, "weights", P.AnimationSampler(
        np.array([0.0, 2.0]), np.array([[0.0, 0.0], [2.0, 4.0]])))])]
    assert np.allclose(first(P.build(s, time=0.5)), base + 0.5 * d0 + 1.0 * d1)  # the animation's win
    assert np.allclose(first(P.build(s)), base + d0 - d1)                        # rest: the node's


def test_skinned_normals_take_the_inverse_transpose():
    """One joint scaled by 2 along X: positions scale, and a normal
    (1, 1, 0)/sqrt(2) becomes (1/2, 1, 0) normalised, the inverse
    transpose's image."""
    r = 1.0 / math.sqrt(2.0)
    prim = P.Primitive(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                       normals=np.array([[r, r, 0]] * 3, np.float32),
                       tangents=np.array([[r, -r, 0, 1]] * 3, np.float32),
                       joints=np.zeros((3, 4), np.uint16), weights=np.array([[1, 0, 0, 0]] * 3, np.float32))
    s = P.Scene3D(meshes=[P.Mesh([prim])], nodes=[P.Node(mesh=0, skin=0), P.Node(scale=(2.0, 1.0, 1.0))],
                  roots=[0, 1], skins=[P.Skin(joints=[1])])
    it = P.build(s).items[0]
    assert it.skinned
    assert np.allclose(it.positions, [[0, 0, 0], [2, 0, 0], [0, 1, 0]])
    n = np.array([0.5, 1.0, 0.0]) / math.sqrt(1.25)
    assert np.allclose(it.normals, n, atol=1e-6)
    # Without influences the primitive is placed by its node, not skinned.
    s.meshes[0].primitives[0].weights = None
    it2 = P.build(s).items[0]
    assert not it2.skinned and np.allclose(it2.positions, [[0, 0, 0], [1, 0, 0], [0, 1, 0]])


def test_mirrored_world_flips_tangent_handedness():
    """B = w (N x T) must transform as M B: under scale (-1, 1, 1) a tangent
    (1, 0, 0, +1) with normal +Z (bitangent +Y) becomes (-1, 0, 0, -1)."""
    prim = P.Primitive(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32),
                       normals=np.array([[0, 0, 1]] * 3, np.float32),
                       tangents=np.array([[1, 0, 0, 1]] * 3, np.float32))
    s = P.Scene3D(meshes=[P.Mesh([prim])], nodes=[P.Node(mesh=0, scale=(-1.0, 1.0, 1.0))], roots=[0])
    it = P.build(s).items[0]
    for k in range(3):
        t, n = it.tangents[k], it.normals[k]
        assert list(t) == [-1.0, 0.0, 0.0, -1.0]
        assert np.allclose(t[3] * np.cross(n, t[:3]), [0.0, 1.0, 0.0])


def test_material_variants():
    prim = P.Primitive(positions=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), material=0,
                       variants=[(1, [0, 2]), (2, [1]), (7, [3])])
    mats = [P.Material(base_color=(c, c, c, 1.0)) for c in (0.1, 0.2, 0.3)]
    s = P.Scene3D(meshes=[P.Mesh([prim])], nodes=[P.Node(mesh=0)], roots=[0], materials=mats)
    got = {v: P.build(s, material_variant=v).items[0].material for v in (None, 0, 1, 2, 3, 9)}
    # Variant 3 maps to a material that does not exist: the fallback.
    assert got == {None: 0, 0: 1, 1: 2, 2: 1, 3: 3, 9: 0}


def test_compute_tangents_vertex_averaged_closed_form():
    """Lengyel's vertex-averaged tangents: a vertex shared by two triangles
    whose dP/du are (1, 0, 0) and (1, 1, 0) gets their normalised sum
    (2, 1, 0)/sqrt(5) (per-face tangents would give each corner its own
    triangle's); flipping v flips w; a tilted normal makes the tangent
    orthogonal to it by Gram-Schmidt."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0]], np.float32)
    uv = np.array([[0, 0], [1, 0], [0, 1], [-1, 1]], np.float32)
    z = np.array([[0, 0, 1]] * 4, np.float32)
    prim = P.Primitive(positions=pos, normals=z, uvs=[uv], indices=np.array([0, 1, 2, 0, 2, 3], np.uint32))
    t = P.compute_tangents(prim, 0)
    r5, r2 = 1 / math.sqrt(5), 1 / math.sqrt(2)
    want = [[2 * r5, r5, 0, 1], [1, 0, 0, 1], [2 * r5, r5, 0, 1], [r2, r2, 0, 1]]
    assert np.allclose(t, want, atol=1e-6)
    flipped = P.Primitive(positions=pos, normals=z, uvs=[uv * np.array([1, -1], np.float32)],
                          indices=prim.indices)
    assert np.allclose(P.compute_tangents(flipped, 0)[:, 3], -1.0)
    tilt = np.array([[0.6, 0, 0.8]] * 4, np.float32)
    one = P.Primitive(positions=pos[:3], normals=tilt[:3], uvs=[uv[:3]])
    assert np.allclose(P.compute_tangents(one, 0), [[0.8, 0, -0.6, 1]] * 3, atol=1e-6)
    # Without vertex normals, or without the UV set, there is nothing to average.
    assert P.compute_tangents(P.Primitive(positions=pos, uvs=[uv]), 0) is None
    assert P.compute_tangents(prim, 1) is None
    # flatten_primitive takes them for a normal-mapped primitive with normals.
    it = P.flatten_primitive(prim, 0, 0, 0, np.eye(4, dtype=np.float32), 0, True, 0)
    assert np.allclose(it.tangents[0], want[0], atol=1e-6) and np.allclose(it.tangents[3], want[0], atol=1e-6)
