"""The prepare layer for the HIP tracer's Python side (oxideav-render
src/prepare.rs): a glTF-shaped scene graph (Scene3D) flattened into the
world-space triangle soup, material table, textures and lights that
`oxr_pt_upload_scene` takes, plus the scene's camera instances and the
camera that Camera::resolve picks among them (src/camera.rs:177-475).

It restates PreparedScene::build (src/prepare.rs:1098-1237): the pose of
an animation at a time (or the rest pose), morph weights by the precedence
animation > node > mesh, skinned primitives placed by their joint palette
in world space, KHR_materials_variants, preorder node traversal from the
roots (`preorder_nodes`, :1282-1296), world transforms,
`flatten_primitive` (:1352-1517: index
streams, strips and fans, garbage indices dropped, winding flipped under a
mirroring world transform, per-corner normals or flat ones, UV sets,
tangents, COLOR_0), `face_tangents` (:1522-1575) for normal-mapped
primitives without tangents, `PreparedMaterial::from_material` /
`fallback` (:599-672), `PreparedLight::from_scene` / `from_light_spec`
(:848-919) and the fallback-light rule. Line and point primitives are
flattened (they count as items) but trace nothing, as in TraceScene.

The scene-graph types mirror oxideav-mesh3d's Scene3D, which is a
crates.io dependency not vendored with the reference, so what that crate
evaluates is restated from the glTF 2.0 specification:
- node transforms compose as T * R * S with (x, y, z, w) quaternions;
- primitives transform with the world matrix (positions), its inverse
  transpose (normals) and its upper 3x3 (tangents, whose handedness w
  flips under a mirroring matrix: B = w (N x T) must map to M B);
- strips and fans triangulate as section 3.7.2.1 defines;
- animation samplers follow Appendix C (STEP, LINEAR with slerp for
  rotations, CUBICSPLINE Hermite splines), clamped outside the keys;
- morph targets add weighted POSITION / NORMAL / TANGENT deltas (3.7.2.2);
- skinning is linear blending of the joint matrices world(joint) x
  inverse bind matrix, which puts the vertices in world space; the mesh
  node's own transform is then ignored (3.7.3.2);
- a variant picks the material of the first KHR_materials_variants
  mapping that lists it.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi as A
from . import scenes as S

F32 = np.float32

TOPOLOGIES = ("triangles", "triangle_strip", "triangle_fan", "lines", "line_strip", "line_loop", "points")


@dataclass
class TextureRef:
    """A material's texture reference: texture index, TEXCOORD set and an
    optional KHR_texture_transform (offset, rotation, scale, and a
    texCoord override)."""
    texture: int
    tex_coord: int = 0
    offset: Optional[tuple] = None
    rotation: float = 0.0
    scale: Optional[tuple] = None
    transform_tex_coord: Optional[int] = None

    def effective_uv_set(self):
        return self.transform_tex_coord if self.transform_tex_coord is not None else self.tex_coord

    def has_transform(self):
        off = self.offset or (0.0, 0.0)
        sc = self.scale or (1.0, 1.0)
        ident = off == (0.0, 0.0) and self.rotation == 0.0 and tuple(sc) == (1.0, 1.0)
        finite = all(math.isfinite(v) for v in (*off, self.rotation, *sc))
        return finite and not ident


@dataclass
class Texture:
    """A decoded image (H, W, 4) uint8 or float32, row 0 at v = 0, with its
    glTF sampler."""
    pixels: np.ndarray
    mag_filter: int = A.FILTER_LINEAR
    min_filter: int = A.MIN_LINEAR_MIP_LINEAR
    wrap_s: int = A.WRAP_REPEAT
    wrap_t: int = A.WRAP_REPEAT


@dataclass
class Material:
    """A glTF metallic-roughness material. `params` holds the remaining
    oxr_material fields (KHR extension scalars: specular, transmission,
    clearcoat, ...); `textures` maps oxr_material_textures slot names to
    TextureRefs."""
    base_color: tuple = (1.0, 1.0, 1.0, 1.0)
    metallic: float = 1.0
    roughness: float = 1.0
    emissive_factor: tuple = (0.0, 0.0, 0.0)
    emissive_strength: Optional[float] = None
    alpha_mode: int = A.OXR_ALPHA_OPAQUE
    alpha_cutoff: float = 0.5
    double_sided: bool = False
    unlit: bool = False
    ior: Optional[float] = None
    normal_scale: float = 1.0
    params: dict = field(default_factory=dict)
    textures: dict = field(default_factory=dict)


@dataclass
class MorphTarget:
    """Per-vertex deltas of one morph target (POSITION, NORMAL, TANGENT
    xyz); each is optional."""
    positions: Optional[np.ndarray] = None
    normals: Optional[np.ndarray] = None
    tangents: Optional[np.ndarray] = None


@dataclass
class Primitive:
    """`joints` / `weights`: JOINTS_n / WEIGHTS_n side by side, (n, 4k)
    each. `variants`: KHR_materials_variants mappings, (material, [variant
    indices]) in the file's order."""
    positions: np.ndarray
    topology: str = "triangles"
    normals: Optional[np.ndarray] = None
    tangents: Optional[np.ndarray] = None
    uvs: List[np.ndarray] = field(default_factory=list)
    colors: List[np.ndarray] = field(default_factory=list)
    indices: Optional[np.ndarray] = None
    material: Optional[int] = None
    targets: List[MorphTarget] = field(default_factory=list)
    joints: Optional[np.ndarray] = None
    weights: Optional[np.ndarray] = None
    variants: List[tuple] = field(default_factory=list)


@dataclass
class Mesh:
    primitives: List[Primitive]
    weights: List[float] = field(default_factory=list)   # default morph weights


@dataclass
class Light:
    """KHR_lights_punctual: kind 'directional', 'point' or 'spot'."""
    kind: str
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    range: Optional[float] = None
    inner_cone_angle: float = 0.0
    outer_cone_angle: float = math.pi / 4.0


@dataclass
class Camera:
    """A glTF camera: perspective (yfov, aspect) or orthographic (xmag,
    ymag)."""
    perspective: bool = True
    yfov: float = 0.8
    aspect: Optional[float] = None
    znear: float = 0.1
    xmag: float = 1.0
    ymag: float = 1.0


@dataclass
class Node:
    mesh: Optional[int] = None
    light: Optional[int] = None
    camera: Optional[int] = None
    children: List[int] = field(default_factory=list)
    matrix: Optional[np.ndarray] = None        # 4x4 row-major (column vectors)
    translation: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)     # quaternion x, y, z, w
    scale: tuple = (1.0, 1.0, 1.0)
    skin: Optional[int] = None
    weights: List[float] = field(default_factory=list)   # morph weights overriding the mesh's


@dataclass
class Skin:
    """Joint nodes and their inverse bind matrices (4x4, column vectors;
    identity when absent)."""
    joints: List[int]
    inverse_bind_matrices: Optional[list] = None
    skeleton: Optional[int] = None


@dataclass
class AnimationSampler:
    """Keyframe times and outputs, one row per key (three per key for
    CUBICSPLINE: in-tangent, value, out-tangent). interpolation: 'STEP',
    'LINEAR' or 'CUBICSPLINE'."""
    times: np.ndarray
    values: np.ndarray
    interpolation: str = "LINEAR"


@dataclass
class Channel:
    node: int
    path: str                    # 'translation', 'rotation', 'scale' or 'weights'
    sampler: AnimationSampler


@dataclass
class Animation:
    channels: List[Channel] = field(default_factory=list)
    name: Optional[str] = None


@dataclass
class Pose:
    """Per node: the animated translation / rotation / scale / morph
    weights (nodes and properties no channel drives are absent)."""
    translation: dict = field(default_factory=dict)
    rotation: dict = field(default_factory=dict)
    scale: dict = field(default_factory=dict)
    weights: dict = field(default_factory=dict)


@dataclass
class Scene3D:
    nodes: List[Node] = field(default_factory=list)
    roots: List[int] = field(default_factory=list)
    meshes: List[Mesh] = field(default_factory=list)
    materials: List[Material] = field(default_factory=list)
    textures: List[Texture] = field(default_factory=list)
    lights: List[Light] = field(default_factory=list)
    cameras: List[Camera] = field(default_factory=list)
    skins: List[Skin] = field(default_factory=list)
    animations: List[Animation] = field(default_factory=list)


@dataclass
class LightSpec:
    """RenderOptions' light (src/options.rs:219-262): the fallback
    directional light."""
    azimuth_deg: float = 45.0
    elevation_deg: float = 45.0
    intensity: float = 1.0


@dataclass
class DrawItem:
    node: int
    mesh: int
    primitive: int
    topology: str               # 'triangles', 'lines' or 'points'
    world: np.ndarray
    positions: np.ndarray       # corners, (n, 3)
    normals: np.ndarray
    has_vertex_normals: bool
    tangents: np.ndarray
    uvs: List[np.ndarray]
    colors: np.ndarray
    material: int
    bounds_min: np.ndarray
    bounds_max: np.ndarray
    skinned: bool = False       # placed by its joint palette (world is then the identity)

    def primitive_count(self):
        per = {"triangles": 3, "lines": 2, "points": 1}[self.topology]
        return len(self.positions) // per


@dataclass
class PreparedScene:
    items: List[DrawItem]
    materials: list             # A.Material, the fallback last
    material_textures: dict     # material index -> slot -> TextureRef
    textures: List[Texture]
    lights: list                # A.Light
    cameras: list               # (node, camera index, Camera, world)
    scene_lights: bool
    bounds: Optional[tuple]

    def default_material(self):
        return len(self.materials) - 1


# ---------------------------------------------------------------------------
# Transforms.
# ---------------------------------------------------------------------------
def _quat_matrix(q):
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], np.float64)


def local_matrix(node, pose=None, i=None):
    """T * R * S of the node, with the pose's animated properties in place
    of the node's own. A node given by a matrix is not animated (glTF
    allows channels only on T/R/S nodes): its matrix stands."""
    if node.matrix is not None:
        return np.asarray(node.matrix, np.float64).reshape(4, 4)
    t, r, sc = node.translation, node.rotation, node.scale
    if pose is not None:
        t = pose.translation.get(i, t)
        r = pose.rotation.get(i, r)
        sc = pose.scale.get(i, sc)
    m = np.eye(4)
    m[:3, :3] = _quat_matrix(r) * np.asarray(sc, np.float64)[None, :]
    m[:3, 3] = t
    return m


def preorder_nodes(scene):
    """src/prepare.rs:1282-1296."""
    visited = [False] * len(scene.nodes)
    out, stack = [], list(reversed(scene.roots))
    while stack:
        i = stack.pop()
        if i < 0 or i >= len(visited) or visited[i]:
            continue
        visited[i] = True
        out.append(i)
        stack.extend(reversed(scene.nodes[i].children))
    return out


def world_transforms(scene, pose=None):
    """World matrix of every node reachable from the roots (first visit),
    in the rest pose or in `pose`."""
    worlds = [None] * len(scene.nodes)
    stack = [(r, np.eye(4)) for r in reversed(scene.roots)]
    while stack:
        i, parent = stack.pop()
        if i < 0 or i >= len(worlds) or worlds[i] is not None:
            continue
        worlds[i] = parent @ local_matrix(scene.nodes[i], pose, i)
        for c in reversed(scene.nodes[i].children):
            stack.append((c, worlds[i]))
    return [None if w is None else w.astype(F32) for w in worlds]


def det3(m):
    """src/prepare.rs:1314-1318."""
    m = np.asarray(m, F32)
    return F32(m[0, 0] * (m[1, 1] * m[2, 2] - m[1, 2] * m[2, 1]) - m[0, 1] * (m[1, 0] * m[2, 2] - m[1, 2] * m[2, 0])
               + m[0, 2] * (m[1, 0] * m[2, 1] - m[1, 1] * m[2, 0]))


def normal_matrix(m):
    """src/prepare.rs:1320-1346: the inverse transpose (cofactors / det)."""
    m = np.asarray(m, F32)
    d = det3(m)
    if not math.isfinite(d) or abs(d) <= 1.0e-20:
        return np.eye(3, dtype=F32)
    inv = F32(1.0) / d

    def c(r0, c0, r1, c1):
        return m[r0, c0] * m[r1, c1] - m[r0, c1] * m[r1, c0]
    return np.array([[c(1, 1, 2, 2) * inv, -c(1, 0, 2, 2) * inv, c(1, 0, 2, 1) * inv],
                     [-c(0, 1, 2, 2) * inv, c(0, 0, 2, 2) * inv, -c(0, 0, 2, 1) * inv],
                     [c(0, 1, 1, 2) * inv, -c(0, 0, 1, 2) * inv, c(0, 0, 1, 1) * inv]], F32)


def _normalise(v):
    v = np.asarray(v, F32)
    n = np.sqrt((v * v).sum(-1, keepdims=True))
    out = np.where(n < F32(1.1920929e-7), F32(0.0), v / np.where(n == 0, F32(1.0), n))
    return out.astype(F32)


def _with(prim, **kw):
    d = dict(prim.__dict__)
    d.update(kw)
    return Primitive(**d)


def transformed(prim, world):
    """The primitive in world space: positions by the world matrix, normals
    by its inverse transpose, tangent directions by its upper 3x3, and the
    tangent handedness flipped when the matrix mirrors."""
    w = np.asarray(world, F32)
    p = np.asarray(prim.positions, F32).reshape(-1, 3)
    pos = (p @ w[:3, :3].T + w[:3, 3]).astype(F32)
    nrm = None
    if prim.normals is not None:
        nrm = (np.asarray(prim.normals, F32).reshape(-1, 3) @ normal_matrix(w).T).astype(F32)
    tan = None
    if prim.tangents is not None:
        t = np.asarray(prim.tangents, F32).reshape(-1, 4)
        hand = F32(-1.0) if det3(w) < 0.0 else F32(1.0)
        tan = np.concatenate([t[:, :3] @ w[:3, :3].T, t[:, 3:4] * hand], 1).astype(F32)
    return _with(prim, positions=pos, normals=nrm, tangents=tan)


# ---------------------------------------------------------------------------
# Animation, morph targets, skinning, variants (glTF 2.0: Appendix C,
# sections 3.7.2.2, 3.7.3 and 3.11; KHR_materials_variants).
# ---------------------------------------------------------------------------
def _slerp(a, b, u):
    """Appendix C: spherical interpolation along the shorter arc,
    sin((1-u) g)/sin g a + s sin(u g)/sin g b with cos g = |a.b| and
    s = sign(a.b); normalised linear interpolation when g is tiny."""
    d = float(np.dot(a, b))
    s = -1.0 if d < 0.0 else 1.0
    ad = min(abs(d), 1.0)
    if ad > 1.0 - 1.0e-6:
        q = (1.0 - u) * a + s * u * b
    else:
        g = math.acos(ad)
        q = (math.sin((1.0 - u) * g) * a + s * math.sin(u * g) * b) / math.sin(g)
    n = math.sqrt(float(np.dot(q, q)))
    return q / n if n > 0.0 else a


def sample_channel(sampler, t, rotation=False):
    """The output of an animation sampler at time t (Appendix C), as a
    float64 vector, or None when the sampler is malformed (no keys, or
    outputs that are not a whole number of rows per key). Before the first
    key and after the last the end values hold; keys that do not increase
    hold the earlier value. Rotations come out normalised."""
    times = np.asarray(sampler.times, np.float64).ravel()
    n = len(times)
    cubic = sampler.interpolation == "CUBICSPLINE"
    vals = np.asarray(sampler.values, np.float64)
    rows = n * (3 if cubic else 1)
    if n == 0 or vals.size == 0 or vals.size % rows:
        return None
    vals = vals.reshape(rows, -1)

    def value(k):
        return vals[3 * k + 1] if cubic else vals[k]

    def out(v):
        if rotation:
            nn = math.sqrt(float(np.dot(v, v)))
            return v / nn if nn > 0.0 else np.array([0.0, 0.0, 0.0, 1.0])
        return v
    if not (t > times[0]):
        return out(value(0).copy())
    if t >= times[-1]:
        return out(value(n - 1).copy())
    k = int(np.searchsorted(times, t, side="right")) - 1
    td = times[k + 1] - times[k]
    if not (td > 0.0) or sampler.interpolation == "STEP":
        return out(value(k).copy())
    u = (t - times[k]) / td
    if cubic:
        u2, u3 = u * u, u * u * u
        b_k, a_k1 = vals[3 * k + 2], vals[3 * (k + 1)]
        v = ((2 * u3 - 3 * u2 + 1) * value(k) + td * (u3 - 2 * u2 + u) * b_k + (-2 * u3 + 3 * u2) * value(k + 1)
             + td * (u3 - u2) * a_k1)
        return out(v)
    if rotation:
        return _slerp(value(k), value(k + 1), u)
    return (1.0 - u) * value(k) + u * value(k + 1)


def sample_pose(anim, t, n_nodes):
    """Animation::sample_pose: every channel on a node in range, sampled at
    t; a channel whose output has the wrong width for its property is
    skipped, and of two channels on one property the later wins."""
    pose = Pose()
    width = {"translation": 3, "rotation": 4, "scale": 3}
    for ch in anim.channels:
        if not (0 <= ch.node < n_nodes):
            continue
        v = sample_channel(ch.sampler, t, rotation=ch.path == "rotation")
        if v is None:
            continue
        if ch.path == "weights":
            pose.weights[ch.node] = [float(x) for x in v]
        elif ch.path in width and len(v) == width[ch.path]:
            getattr(pose, ch.path)[ch.node] = tuple(float(x) for x in v)
    return pose


def morphed(prim, weights):
    """Primitive::morphed: base attributes plus the weighted deltas of each
    target (a missing or non-finite weight counts as 0; deltas of the wrong
    length are skipped; tangent deltas move xyz only)."""
    if not prim.targets or not len(weights):
        return prim
    pos = np.array(prim.positions, F32).reshape(-1, 3)
    n = len(pos)
    nrm = None if prim.normals is None else np.array(prim.normals, F32).reshape(-1, 3)
    tan = None if prim.tangents is None else np.array(prim.tangents, F32).reshape(-1, 4)
    for w, tg in zip(weights, prim.targets):
        w = F32(w) if math.isfinite(w) else F32(0.0)
        if w == 0.0:
            continue
        if tg.positions is not None and len(tg.positions) == n:
            pos += w * np.asarray(tg.positions, F32).reshape(-1, 3)
        if nrm is not None and tg.normals is not None and len(tg.normals) == n:
            nrm += w * np.asarray(tg.normals, F32).reshape(-1, 3)
        if tan is not None and tg.tangents is not None and len(tg.tangents) == n:
            tan[:, :3] += w * np.asarray(tg.tangents, F32).reshape(n, -1)[:, :3]
    return _with(prim, positions=pos, normals=nrm, tangents=tan)


def joint_matrices(scene, nid, worlds):
    """Scene3D::joint_matrices_with: world(joint j) x inverse bind matrix j
    for the node's skin, or None when the node has no valid skin (a joint
    outside the node list). A joint that the roots do not reach takes its
    local matrix as its world matrix."""
    node = scene.nodes[nid]
    if node.skin is None or not (0 <= node.skin < len(scene.skins)):
        return None
    skin = scene.skins[node.skin]
    ibms = skin.inverse_bind_matrices or []
    pal = []
    for j, jn in enumerate(skin.joints):
        if not (0 <= jn < len(scene.nodes)):
            return None
        w = worlds[jn] if worlds[jn] is not None else local_matrix(scene.nodes[jn])
        ibm = np.asarray(ibms[j], np.float64).reshape(4, 4) if j < len(ibms) else np.eye(4)
        pal.append(np.asarray(w, np.float64) @ ibm)
    return np.asarray(pal, F32).reshape(-1, 4, 4) if pal else None


def has_influences(prim):
    n = len(np.asarray(prim.positions).reshape(-1, 3))
    return (prim.joints is not None and prim.weights is not None
            and np.asarray(prim.joints).reshape(n, -1).shape == np.asarray(prim.weights).reshape(n, -1).shape)


def skinned(prim, palette):
    """Primitive::skinned: each vertex by the weighted sum of its joints'
    palette matrices (linear blend skinning) into world space; normals by
    that sum's inverse transpose, tangents by its upper 3x3 with the
    handedness flipped where it mirrors. Influences naming a joint outside
    the palette, or with a non-finite weight, count as 0."""
    p = np.asarray(prim.positions, F32).reshape(-1, 3)
    n = len(p)
    J = np.asarray(prim.joints).reshape(n, -1).astype(np.int64)
    W = np.asarray(prim.weights, F32).reshape(n, -1)
    ok = (J >= 0) & (J < len(palette)) & np.isfinite(W)
    Wv = np.where(ok, W, F32(0.0)).astype(F32)
    M = np.einsum("nk,nkij->nij", Wv, palette[np.where(ok, J, 0)]).astype(F32)
    R = M[:, :3, :3]
    pos = (np.einsum("nij,nj->ni", R, p) + M[:, :3, 3]).astype(F32)
    det, nm = _normal_matrices(R)
    nrm = None
    if prim.normals is not None:
        nv = np.asarray(prim.normals, F32).reshape(-1, 3)
        nrm = np.einsum("nij,nj->ni", nm, nv).astype(F32)
    tan = None
    if prim.tangents is not None:
        t = np.asarray(prim.tangents, F32).reshape(-1, 4)
        hand = np.where(det < 0.0, F32(-1.0), F32(1.0)).astype(F32).reshape(-1, 1)
        tan = np.concatenate([np.einsum("nij,nj->ni", R, t[:, :3]), t[:, 3:4] * hand], 1).astype(F32)
    return _with(prim, positions=pos, normals=nrm, tangents=tan)


def _normal_matrices(R):
    """det3 and normal_matrix of each 3x3 in R (n, 3, 3), vectorised with
    the same f32 operations."""
    R = np.asarray(R, F32)
    m = lambda i, j: R[:, i, j]
    c = lambda r0, c0, r1, c1: m(r0, c0) * m(r1, c1) - m(r0, c1) * m(r1, c0)
    det = (m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) - m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0))
           + m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0))).astype(F32)
    ok = np.isfinite(det) & (np.abs(det) > 1.0e-20)
    inv = np.where(ok, F32(1.0) / np.where(ok, det, F32(1.0)), F32(0.0)).astype(F32)
    nm = np.stack([np.stack([c(1, 1, 2, 2) * inv, -c(1, 0, 2, 2) * inv, c(1, 0, 2, 1) * inv], -1),
                   np.stack([-c(0, 1, 2, 2) * inv, c(0, 0, 2, 2) * inv, -c(0, 0, 2, 1) * inv], -1),
                   np.stack([c(0, 1, 1, 2) * inv, -c(0, 0, 1, 2) * inv, c(0, 0, 1, 1) * inv], -1)], 1).astype(F32)
    nm[~ok] = np.eye(3, dtype=F32)
    return det, nm


def material_for_variant(prim, variant):
    """KHR_materials_variants: the material of the first mapping that lists
    `variant`, else the primitive's own."""
    if variant is not None:
        for mat, vs in prim.variants:
            if variant in vs:
                return mat
    return prim.material


def triangle_indices(prim, seq):
    """Triangles of a triangle, strip or fan index stream (glTF 2.0
    section 3.7.2.1)."""
    n = len(seq)
    if prim.topology == "triangles":
        return [tuple(seq[3 * i:3 * i + 3]) for i in range(n // 3)]
    if prim.topology == "triangle_strip":
        return [(seq[i], seq[i + 1 + i % 2], seq[i + 2 - i % 2]) for i in range(max(n - 2, 0))]
    if prim.topology == "triangle_fan":
        return [(seq[i + 1], seq[i + 2], seq[0]) for i in range(max(n - 2, 0))]
    return []


def face_tangents(pos, nrm, uv):
    """src/prepare.rs:1522-1575 (Lengyel 2001, Gram-Schmidt per corner)."""
    out = np.zeros((len(pos), 4), F32)
    for t in range(len(pos) // 3):
        p0, p1, p2 = pos[3 * t], pos[3 * t + 1], pos[3 * t + 2]
        w0, w1, w2 = uv[3 * t], uv[3 * t + 1], uv[3 * t + 2]
        e1, e2 = p1 - p0, p2 - p0
        du1, dv1 = w1[0] - w0[0], w1[1] - w0[1]
        du2, dv2 = w2[0] - w0[0], w2[1] - w0[1]
        det = F32(du1 * dv2 - du2 * dv1)
        if abs(det) > 1.0e-12:
            r = F32(1.0) / det
            tan = ((e1 * dv2 - e2 * dv1) * r).astype(F32)
            bit = ((e2 * du1 - e1 * du2) * r).astype(F32)
        else:
            tan, bit = np.array([1, 0, 0], F32), np.array([0, 1, 0], F32)
        for k in range(3):
            n = nrm[3 * t + k] if 3 * t + k < len(nrm) else np.array([0, 0, 1], F32)
            d = F32(np.dot(n, tan))
            tt = _normalise(tan - n * d)
            if np.dot(tt, tt) < 0.5:
                a = np.array([1, 0, 0], F32) if abs(n[0]) < 0.9 else np.array([0, 1, 0], F32)
                tt = _normalise(np.cross(a, n))
            w = -1.0 if np.dot(np.cross(n, tt), bit) < 0.0 else 1.0
            out[3 * t + k] = (*tt, w)
    return out


def _perpendicular(n):
    a = np.where((np.abs(n[:, 0]) < 0.9)[:, None], np.array([1, 0, 0], F32), np.array([0, 1, 0], F32))
    return _normalise(np.cross(a, n))


def compute_tangents(prim, uv_set):
    """Primitive::compute_tangents (oxideav-mesh3d; the reference's notes
    call it vertex-averaged per-triangle UV gradients, src/prepare.rs:
    29-32), after Lengyel, "Computing Tangent Space Basis Vectors for an
    Arbitrary Mesh" (2001): each triangle's dP/du and dP/dv are summed
    into its three vertices; a vertex's tangent is its dP/du sum made
    orthogonal to the vertex normal and normalised, and w is the sign of
    (N x T) . (dP/dv sum). Triangles with a degenerate UV map add nothing;
    a vertex that gets nothing (or a sum parallel to its normal) takes a
    perpendicular of the normal, w = +1. None without vertex normals or
    that UV set."""
    p = np.asarray(prim.positions, F32).reshape(-1, 3)
    nv = len(p)
    if prim.normals is None or len(prim.normals) != nv or uv_set >= len(prim.uvs) or len(prim.uvs[uv_set]) != nv:
        return None
    nrm = np.asarray(prim.normals, F32).reshape(-1, 3)
    uv = np.asarray(prim.uvs[uv_set], F32).reshape(-1, 2)
    seq = list(range(nv)) if prim.indices is None else [int(i) for i in np.asarray(prim.indices).ravel()]
    tri = np.array([t for t in triangle_indices(prim, seq) if all(0 <= i < nv for i in t)], np.int64).reshape(-1, 3)
    tan1 = np.zeros((nv, 3), F32)
    tan2 = np.zeros((nv, 3), F32)
    if len(tri):
        e1, e2 = p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]
        d1, d2 = uv[tri[:, 1]] - uv[tri[:, 0]], uv[tri[:, 2]] - uv[tri[:, 0]]
        det = (d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]).astype(F32)
        good = np.abs(det) > 1.0e-12
        r = np.where(good, F32(1.0) / np.where(good, det, F32(1.0)), F32(0.0)).astype(F32)[:, None]
        sdir = ((e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) * r).astype(F32)
        tdir = ((e2 * d1[:, 0:1] - e1 * d2[:, 0:1]) * r).astype(F32)
        for k in range(3):
            np.add.at(tan1, tri[:, k], sdir)
            np.add.at(tan2, tri[:, k], tdir)
    t = _normalise(tan1 - nrm * (nrm * tan1).sum(1, keepdims=True))
    bad = (t * t).sum(1) < 0.5
    t[bad] = _perpendicular(nrm[bad])
    w = np.where((np.cross(nrm, t) * tan2).sum(1) < 0.0, F32(-1.0), F32(1.0))
    return np.concatenate([t, w[:, None]], 1).astype(F32)


def flatten_primitive(prim, node, mesh, pi, world, material, needs_tangents, tangent_uv_set):
    """src/prepare.rs:1352-1517."""
    pos_all = np.asarray(prim.positions, F32).reshape(-1, 3)
    nv = len(pos_all)
    if nv == 0:
        return None
    normals_src = None if prim.normals is None or len(prim.normals) != nv else np.asarray(prim.normals, F32)
    tangents_src = None if prim.tangents is None or len(prim.tangents) != nv else np.asarray(prim.tangents, F32)
    topo = ("triangles" if prim.topology in TOPOLOGIES[:3] else "lines" if prim.topology in TOPOLOGIES[3:6]
            else "points")
    if tangents_src is None and needs_tangents and topo == "triangles" and normals_src is not None:
        tangents_src = compute_tangents(prim, tangent_uv_set)
    seq = list(range(nv)) if prim.indices is None else [int(i) for i in np.asarray(prim.indices).ravel()]
    flip = det3(world) < 0.0
    finite = np.all(np.isfinite(pos_all), axis=1)

    def ok(i):
        return 0 <= i < nv and bool(finite[i])
    corners = []
    if topo == "triangles":
        for a, b, c in triangle_indices(prim, seq):
            if ok(a) and ok(b) and ok(c):
                corners.extend((a, c, b) if flip else (a, b, c))
    elif topo == "lines":
        if prim.topology == "lines":
            pairs = [(seq[2 * i], seq[2 * i + 1]) for i in range(len(seq) // 2)]
        else:
            pairs = [(seq[i], seq[i + 1]) for i in range(max(len(seq) - 1, 0))]
            if prim.topology == "line_loop" and len(seq) >= 2:
                pairs.append((seq[-1], seq[0]))
        for a, b in pairs:
            if ok(a) and ok(b):
                corners.extend((a, b))
    else:
        corners = [i for i in seq if ok(i)]
    if not corners:
        return None
    idx = np.asarray(corners, np.int64)
    positions = pos_all[idx]
    if normals_src is not None:
        nrm = _normalise(normals_src[idx])
        nrm = np.where(np.all(np.isfinite(nrm), axis=1, keepdims=True), nrm, np.array([0, 0, 1], F32)).astype(F32)
        has_vn = True
    elif topo == "triangles":
        t = positions.reshape(-1, 3, 3)
        fn = _normalise(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]))
        nrm = np.repeat(fn, 3, axis=0).astype(F32)
        has_vn = False
    else:
        nrm, has_vn = np.zeros((0, 3), F32), False
    uvs = [np.asarray(u, F32)[idx] if len(u) == nv else np.zeros((0, 2), F32) for u in prim.uvs]
    if tangents_src is not None:
        d = _normalise(tangents_src[idx, :3])
        tangents = np.concatenate([d, np.where(tangents_src[idx, 3:4] < 0.0, -1.0, 1.0)], 1).astype(F32)
    else:
        tangents = np.zeros((0, 4), F32)
    if (len(tangents) == 0 and needs_tangents and topo == "triangles" and tangent_uv_set < len(uvs)
            and len(uvs[tangent_uv_set])):
        tangents = face_tangents(positions, nrm, uvs[tangent_uv_set])
    colors = (np.asarray(prim.colors[0], F32)[idx] if prim.colors and len(prim.colors[0]) == nv
              else np.zeros((0, 4), F32))
    return DrawItem(node=node, mesh=mesh, primitive=pi, topology=topo, world=np.asarray(world, F32),
                    positions=positions, normals=nrm, has_vertex_normals=has_vn, tangents=tangents, uvs=uvs,
                    colors=colors, material=material, bounds_min=positions.min(0), bounds_max=positions.max(0))


# ---------------------------------------------------------------------------
# Materials and lights.
# ---------------------------------------------------------------------------
def _fin(v, d):
    return v if math.isfinite(v) else d


def prepared_material(m):
    """PreparedMaterial::from_material's factor rules (src/prepare.rs:
    623-668) as an oxr_material (the library applies material_oriented's
    finite/clamp rules to the rest at upload)."""
    strength = max(_fin(m.emissive_strength, 1.0), 0.0) if m.emissive_strength is not None else 1.0
    base = tuple(_fin(v, 1.0) for v in m.base_color[:3]) + (min(max(_fin(m.base_color[3], 1.0), 0.0), 1.0),)
    out = A.material(base, metallic=min(max(_fin(m.metallic, 1.0), 0.0), 1.0),
                     roughness=min(max(_fin(m.roughness, 1.0), 0.0), 1.0), **m.params)
    out.emissive = A.f3(*(max(_fin(v, 0.0), 0.0) * strength for v in m.emissive_factor))
    out.flags = (A.OXR_MAT_DOUBLE_SIDED if m.double_sided else 0) | (A.OXR_MAT_UNLIT if m.unlit else 0)
    out.alpha_mode = m.alpha_mode
    out.alpha_cutoff = m.alpha_cutoff
    out.ior = max(_fin(m.ior, 1.5), 1.0) if m.ior is not None else 1.5
    return out


def fallback_material():
    """PreparedMaterial::fallback (src/prepare.rs:603-621)."""
    return A.material((0.7, 0.7, 0.75, 1.0), metallic=0.0, roughness=1.0)


def light_from_spec(spec):
    """PreparedLight::from_light_spec (src/prepare.rs:848-863)."""
    az, el = math.radians(spec.azimuth_deg), math.radians(spec.elevation_deg)
    toward = _normalise(np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)], F32))
    return A.light(A.OXR_LIGHT_DIRECTIONAL, color=(1.0, 1.0, 1.0), intensity=math.pi * max(spec.intensity, 0.0),
                   position=(0.0, 0.0, 0.0), direction=tuple(float(-v) for v in toward), range_=0.0, inner=0.0,
                   outer=math.pi / 4.0)


def light_from_scene(light, world):
    """PreparedLight::from_scene (src/prepare.rs:865-919)."""
    w = np.asarray(world, F32)
    position = (float(w[0, 3]), float(w[1, 3]), float(w[2, 3]))
    direction = _normalise(np.array([-w[0, 2], -w[1, 2], -w[2, 2]], F32))
    if np.dot(direction, direction) < 0.5:
        direction = np.array([0.0, 0.0, -1.0], F32)
    kind = {"directional": A.OXR_LIGHT_DIRECTIONAL, "point": A.OXR_LIGHT_POINT, "spot": A.OXR_LIGHT_SPOT}[light.kind]
    rng = light.range if (light.range is not None and math.isfinite(light.range) and light.range > 0.0) else 0.0
    inner = outer = 0.0
    if light.kind == "directional":
        rng = 0.0
    elif light.kind == "spot":
        oc, ic = light.outer_cone_angle, light.inner_cone_angle
        outer = min(max(oc, 1.0e-4), math.pi / 2.0) if math.isfinite(oc) else math.pi / 4.0
        inner = min(max(ic, 0.0), outer) if math.isfinite(ic) else 0.0
    return A.light(kind, color=tuple(max(_fin(c, 0.0), 0.0) for c in light.color),
                   intensity=max(_fin(light.intensity, 0.0), 0.0), position=position,
                   direction=tuple(float(v) for v in direction), range_=rng, inner=inner, outer=outer)


# ---------------------------------------------------------------------------
# PreparedScene::build.
# ---------------------------------------------------------------------------
def build(scene, use_scene_lights=True, fallback_light=LightSpec(), generate_tangents=True, time=None,
          animation=None, material_variant=None):
    """PreparedScene::build (src/prepare.rs:1098-1237). With a finite
    `time`, animation `animation` (default 0; none when out of range) is
    sampled into the pose the scene is built in; otherwise the rest pose.
    `material_variant` selects a KHR_materials_variants variant."""
    ntex = len(scene.textures)
    materials, mat_tex = [], {}
    for i, m in enumerate(scene.materials):
        materials.append(prepared_material(m))
        binds = {slot: r for slot, r in m.textures.items() if 0 <= r.texture < ntex}
        if binds:
            mat_tex[i] = (binds, m)
    default_material = len(materials)
    materials.append(fallback_material())
    pose = None
    if time is not None and math.isfinite(time):
        ai = 0 if animation is None else animation
        if 0 <= ai < len(scene.animations):
            pose = sample_pose(scene.animations[ai], float(time), len(scene.nodes))
    worlds = world_transforms(scene, pose)
    items, lights, cameras = [], [], []
    for nid in preorder_nodes(scene):
        node, world = scene.nodes[nid], worlds[nid]
        if world is None:
            continue
        if node.light is not None and 0 <= node.light < len(scene.lights):
            lights.append(light_from_scene(scene.lights[node.light], world))
        if node.camera is not None and 0 <= node.camera < len(scene.cameras):
            cameras.append((nid, node.camera, scene.cameras[node.camera], world))
        if node.mesh is None or not (0 <= node.mesh < len(scene.meshes)):
            continue
        mesh = scene.meshes[node.mesh]
        if pose is not None and nid in pose.weights:
            weights = pose.weights[nid]
        else:
            weights = node.weights if len(node.weights) else mesh.weights
        palette = joint_matrices(scene, nid, worlds) if node.skin is not None else None
        for pi, prim in enumerate(mesh.primitives):
            mv = material_for_variant(prim, material_variant)
            mi = mv if (mv is not None and 0 <= mv < default_material) else default_material
            normal_ref = mat_tex.get(mi, ({}, None))[0].get("normal")
            needs = generate_tangents and normal_ref is not None
            m = morphed(prim, weights)
            if palette is not None and has_influences(m):
                baked, item_world, sk = skinned(m, palette), np.eye(4, dtype=F32), True
            else:
                baked, item_world, sk = transformed(m, world), world, False
            it = flatten_primitive(baked, nid, node.mesh, pi, item_world, mi, needs,
                                   normal_ref.effective_uv_set() if normal_ref else 0)
            if it is not None:
                it.skinned = sk
                items.append(it)
    scene_lights = use_scene_lights and len(lights) > 0
    if not scene_lights:
        lights = [light_from_spec(fallback_light)] if fallback_light is not None else []
    bounds = None
    for it in items:
        bounds = (it.bounds_min, it.bounds_max) if bounds is None else (
            np.minimum(bounds[0], it.bounds_min), np.maximum(bounds[1], it.bounds_max))
    return PreparedScene(items=items, materials=materials, material_textures=mat_tex, textures=scene.textures,
                         lights=lights, cameras=cameras, scene_lights=scene_lights, bounds=bounds)


# ---------------------------------------------------------------------------
# Cameras (src/camera.rs:177-475).
# ---------------------------------------------------------------------------
def _n3(v):
    v = np.asarray(v, F32)
    n = F32(math.sqrt(float(np.dot(v, v))))
    return (v / n).astype(F32) if n >= F32(1.1920929e-7) else np.zeros(3, F32)


def _camera(eye, fwd, side, up, half_w, half_h, projection):
    c = A.Camera()
    c.eye, c.forward, c.side, c.up = (A.f3(*(float(x) for x in v)) for v in (eye, fwd, side, up))
    c.half_w, c.half_h, c.projection = float(half_w), float(half_h), projection
    return c


def camera_from_instance(cam, world, width, height):
    """Camera::from_scene_camera (src/camera.rs:208-332): a glTF camera at
    its node's world matrix (scale ignored), filling the viewport aspect."""
    w = np.asarray(world, F32)
    eye = w[:3, 3].copy()
    fwd = _n3(-w[:3, 2])
    if np.dot(fwd, fwd) < 0.5:
        fwd = np.array([0.0, 0.0, -1.0], F32)
    side = _n3(np.cross(fwd, _n3(w[:3, 1])))
    if np.dot(side, side) < 0.5:
        side = _n3(np.cross(fwd, np.array([0.0, 1.0, 0.0], F32)))
        if np.dot(side, side) < 0.5:
            side = np.array([1.0, 0.0, 0.0], F32)
    up = np.cross(side, fwd).astype(F32)
    aspect = F32(max(width, 1)) / F32(max(height, 1))
    if cam.perspective:
        yfov = cam.yfov if (math.isfinite(cam.yfov) and 1.0e-4 < cam.yfov < math.pi - 1.0e-3) else math.pi / 3.0
        th = F32(math.tan(float(F32(yfov) * F32(0.5))))
        return _camera(eye, fwd, side, up, th * aspect, th, A.OXR_PROJ_PERSPECTIVE)
    hh = F32(abs(cam.ymag)) if (math.isfinite(cam.ymag) and abs(cam.ymag) > 1.0e-6) else F32(1.0)
    return _camera(eye, fwd, side, up, hh * aspect, hh, A.OXR_PROJ_ORTHOGRAPHIC)


def frame_bounds(width, height, bounds, projection=A.OXR_PROJ_PERSPECTIVE, fov_deg=60.0, orbit=None,
                 target_offset=(0.0, 0.0, 0.0)):
    """Camera::build (src/camera.rs:357-475): auto-frame the AABB, looking
    down -Z, or orbit it when `orbit` = (elevation_deg, azimuth_deg,
    distance multiplier) is given (RenderOptions::camera)."""
    mn, mx = (np.asarray(b, F32) for b in bounds)
    off = np.asarray(target_offset, F32)
    if not np.all(np.isfinite(off)):
        off = np.zeros(3, F32)
    c = ((mn + mx) * F32(0.5) + off).astype(F32)
    extent = F32(max(float(mx[0] - mn[0]), float(mx[1] - mn[1]), float(mx[2] - mn[2]), 1.0e-3))
    aspect = F32(max(width, 1)) / F32(max(height, 1))
    radius = extent * F32(0.5) * F32(1.2)
    fov_y = F32(math.radians(fov_deg))
    persp = projection == A.OXR_PROJ_PERSPECTIVE
    auto = radius / F32(math.tan(float(fov_y * F32(0.5)))) if persp else extent * F32(1.5)
    if orbit is None:
        eye = c + np.array([0.0, 0.0, auto], F32)
    else:
        el, az = math.radians(orbit[0]), math.radians(orbit[1])
        d = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)], F32)
        eye = (c + d * F32(auto * F32(orbit[2]))).astype(F32)
    world_up = np.array([0.0, 1.0, 0.0], F32)
    fwd = _n3(c - eye)
    side = _n3(np.cross(fwd, world_up))
    up = np.cross(side, fwd).astype(F32)
    if persp:
        th = F32(math.tan(float(fov_y * F32(0.5))))
        return _camera(eye, fwd, side, up, th * aspect, th, projection)
    zoom = orbit[2] if (orbit is not None and math.isfinite(orbit[2]) and orbit[2] > 0.0) else 1.0
    half = radius * F32(zoom)
    hw, hh = (half * aspect, half) if aspect >= 1.0 else (half, half / aspect)
    return _camera(eye, fwd, side, up, hw, hh, projection)


class PreparedCamera:
    """Camera::resolve (src/camera.rs:177-190) for a PreparedScene: the
    scene camera instance `scene_camera` when it exists, otherwise the
    auto-framed (or orbiting) camera over the prepared bounds (a unit box
    around the origin when there are none, src/prepare.rs:1259-1261).
    `resolve(width, height)` gives the oxr_camera."""

    def __init__(self, prep, scene_camera=None, projection=A.OXR_PROJ_PERSPECTIVE, fov_deg=60.0, orbit=None,
                 target_offset=(0.0, 0.0, 0.0)):
        self.bounds = prep.bounds if prep.bounds is not None else ((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5))
        self.instance = (prep.cameras[scene_camera] if scene_camera is not None
                         and 0 <= scene_camera < len(prep.cameras) else None)
        self.projection, self.fov_deg, self.orbit, self.target_offset = projection, fov_deg, orbit, target_offset

    def resolve(self, width, height):
        if self.instance is not None:
            _, _, cam, world = self.instance
            return camera_from_instance(cam, world, width, height)
        return frame_bounds(width, height, self.bounds, self.projection, self.fov_deg, self.orbit,
                            self.target_offset)


def to_scene(prep, name="prepared", scene_camera=None):
    """The tracer's scene (oxr_scene_desc) of a PreparedScene: the triangle
    items' corners in order, as TraceScene::new takes them
    (src/trace.rs:329-343), and the camera Camera::resolve picks
    (`scene_camera` = RenderOptions::scene_camera; replace `.camera` for
    an orbit or another field of view)."""
    s = S.Scene(name)
    for m in prep.materials:
        s.add_material(m)
    for t in prep.textures:
        s.add_texture(t.pixels, mag=t.mag_filter, min_filter=t.min_filter, wrap_s=t.wrap_s, wrap_t=t.wrap_t)
    for mi, (binds, m) in prep.material_textures.items():
        slots = {}
        for slot, r in binds.items():
            if slot not in A.TEXTURE_SLOTS:
                continue  # occlusion: the path tracer ignores it (src/pathtrace.rs:241-246)
            b = dict(texture=r.texture, uv_set=r.effective_uv_set())
            if r.has_transform():
                b.update(offset=tuple(r.offset or (0.0, 0.0)), rotation=r.rotation, scale=tuple(r.scale or (1.0, 1.0)))
            slots[slot] = b
        extra = {"normal_scale": _fin(m.normal_scale, 1.0)}
        extra.update({k: m.params[k] for k in ("clearcoat_normal_scale", "iridescence_thickness_min")
                      if k in m.params})
        s.set_material_textures(mi, **extra, **slots)
    for it in prep.items:
        if it.topology != "triangles":
            continue
        n = len(it.positions) // 3
        kw = {}
        for k in range(min(2, len(it.uvs))):
            if len(it.uvs[k]):
                kw["uv%d" % k] = it.uvs[k].reshape(n, 6)
        if len(it.tangents):
            kw["tangents"] = it.tangents.reshape(n, 12)
        if len(it.colors):
            kw["colors"] = it.colors.reshape(n, 12)
        s.add_tris(it.positions.reshape(n, 9), it.material,
                   normals=it.normals.reshape(n, 9) if it.has_vertex_normals else None, **kw)
    s.lights = list(prep.lights)
    s.camera = PreparedCamera(prep, scene_camera)
    return s


# ---------------------------------------------------------------------------
# An example scene graph.
# ---------------------------------------------------------------------------
def showcase_graph(res=32):
    """A small glTF-shaped scene graph that exercises the prepare rules: a
    strip floor scaled by its node, an indexed box with COLOR_0 placed twice
    (once under a mirroring node), a normal-mapped textured panel without
    tangents (generated per face), an emissive fan, a line loop and a
    material-less primitive, lit by a spot and a point light on
    transformed nodes and seen through a camera node."""
    u = (np.arange(res) + 0.5) / res
    uu, vv = np.meshgrid(u, u)
    checker = ((np.floor(uu * 4) + np.floor(vv * 4)) % 2)[..., None]
    base = np.concatenate([40 + 180 * checker, 90 + 60 * checker, 200 - 150 * checker,
                           np.full_like(checker, 255)], 2).astype(np.uint8)
    nx, ny = 0.5 * np.sin(uu * 6 * math.pi), 0.5 * np.cos(vv * 6 * math.pi)
    nz = np.ones_like(nx)
    ln = np.sqrt(nx * nx + ny * ny + nz * nz)
    nmap = np.stack([nx / ln, ny / ln, nz / ln], 2) * 127.5 + 127.5
    nmap = np.concatenate([nmap, np.full(nmap.shape[:2] + (1,), 255)], 2).round().astype(np.uint8)
    textures = [Texture(base), Texture(nmap)]
    materials = [
        Material(base_color=(0.8, 0.8, 0.8, 1.0), metallic=0.0, roughness=0.7),
        Material(base_color=(1.0, 1.0, 1.0, 1.0), metallic=0.0, roughness=0.4),
        Material(base_color=(1.0, 1.0, 1.0, 1.0), metallic=0.2, roughness=0.35, double_sided=True,
                 textures={"base_color": TextureRef(0), "normal": TextureRef(1)}, normal_scale=0.8),
        Material(base_color=(0.0, 0.0, 0.0, 1.0), emissive_factor=(1.0, 0.8, 0.6), emissive_strength=6.0),
    ]
    quad = np.array([[-1, 0, 1], [1, 0, 1], [-1, 0, -1], [1, 0, -1]], np.float32)
    floor = Primitive(positions=quad, topology="triangle_strip", normals=np.array([[0, 1, 0]] * 4, np.float32),
                      material=0)
    c = np.array([[x, y, z] for x in (-0.3, 0.3) for y in (0.0, 0.6) for z in (-0.3, 0.3)], np.float32)
    box_idx = np.array([0, 6, 4, 0, 2, 6, 1, 7, 3, 1, 5, 7, 0, 5, 1, 0, 4, 5,
                        2, 7, 6, 2, 3, 7, 0, 3, 2, 0, 1, 3, 4, 7, 5, 4, 6, 7], np.uint16)
    cols = np.concatenate([(c[:, :1] + 0.3) / 0.6, (c[:, 1:2]) / 0.6, (c[:, 2:3] + 0.3) / 0.6,
                           np.ones((8, 1), np.float32)], 1).astype(np.float32)
    box = Primitive(positions=c, indices=box_idx, colors=[cols], material=1)
    panel = Primitive(positions=np.array([[-0.5, 0, 0], [0.5, 0, 0], [0.5, 1, 0], [-0.5, 1, 0]], np.float32),
                      indices=np.array([0, 1, 2, 0, 2, 3], np.uint32),
                      normals=np.array([[0, 0, 1]] * 4, np.float32),
                      uvs=[np.array([[0, 1], [1, 1], [1, 0], [0, 0]], np.float32)], material=2)
    ring = [[0.0, 0.0, 0.0]] + [[0.25 * math.cos(a), 0.0, -0.25 * math.sin(a)]
                                for a in np.linspace(0.0, 2.0 * math.pi, 9)]
    lamp = Primitive(positions=np.array(ring, np.float32), topology="triangle_fan", material=3)
    wire = Primitive(positions=quad, topology="line_loop", material=0)
    plain = Primitive(positions=np.array([[1.2, 0.01, 0.8], [1.6, 0.01, 0.8], [1.4, 0.5, 0.6]], np.float32))
    nodes = [
        Node(children=[1, 2, 3, 4, 5, 6, 7, 8, 9]),
        Node(mesh=0, scale=(3.0, 1.0, 3.0)),                                   # floor
        Node(mesh=1, translation=(-0.9, 0.0, 0.2), rotation=(0.0, 0.2588190, 0.0, 0.9659258)),
        Node(mesh=1, translation=(0.9, 0.0, -0.4), scale=(-1.0, 1.3, 1.0)),   # mirrored box
        Node(mesh=2, translation=(0.0, 0.05, -1.2), rotation=(0.0, 0.0, 0.0, 1.0), scale=(1.6, 1.6, 1.6)),
        Node(mesh=3, translation=(0.3, 2.2, 0.3), rotation=(1.0, 0.0, 0.0, 0.0)),  # fan faces down
        Node(light=0, translation=(0.0, 3.0, 1.5), rotation=(-0.4, 0.0, 0.0, 0.9165151)),
        Node(light=1, translation=(-1.5, 1.5, 1.0)),
        Node(mesh=4),
        Node(camera=0, translation=(0.0, 1.7, 3.4), rotation=(-0.1522881, 0.0, 0.0, 0.9883358)),  # 17.5 deg down
    ]
    meshes = [Mesh([floor, wire]), Mesh([box]), Mesh([panel]), Mesh([lamp]), Mesh([plain])]
    lights = [Light("spot", color=(1.0, 0.9, 0.8), intensity=30.0, range=12.0, inner_cone_angle=0.3,
                    outer_cone_angle=0.6),
              Light("point", color=(0.6, 0.7, 1.0), intensity=4.0)]
    return Scene3D(nodes=nodes, roots=[0], meshes=meshes, materials=materials, textures=textures, lights=lights,
                   cameras=[Camera(yfov=0.9)])


def skinned_morph_beam():
    """The reference's animated test scene (src/testscenes.rs:576-680): a
    2 x 0.2 strip along +X (21 vertex columns) skinned to two joints, root
    at the origin and `elbow` at x = 1, with weights ramping linearly over
    x in [0.5, 1.5], and one morph target lifting every vertex by 0.3 in Y.
    Animation 0 turns the elbow from 0 to 90 degrees about +Z and drives the
    morph weight from 0 to 1 over t in [0, 1]. Nodes: 0 elbow, 1 root, 2 the
    mesh, 3 a directional light along -Z, 4 a camera at (1, 0.5, 4) looking
    down -Z (yfov 0.9)."""
    cols = 21
    pos, joints, weights = [], [], []
    for c in range(cols):
        x = 2.0 * c / (cols - 1)
        w1 = min(max(x - 0.5, 0.0), 1.0)
        for y in (-0.1, 0.1):
            pos.append((x, y, 0.0))
            joints.append((0, 1, 0, 0))
            weights.append((1.0 - w1, w1, 0.0, 0.0))
    idx = []
    for c in range(cols - 1):
        a = 2 * c
        idx += [a, a + 2, a + 3, a, a + 3, a + 1]
    n = len(pos)
    prim = Primitive(positions=np.array(pos, F32), indices=np.array(idx, np.uint32),
                     normals=np.tile(np.array([0.0, 0.0, 1.0], F32), (n, 1)), material=0,
                     targets=[MorphTarget(positions=np.tile(np.array([0.0, 0.3, 0.0], F32), (n, 1)))],
                     joints=np.array(joints, np.uint16), weights=np.array(weights, F32))
    ibm1 = np.eye(4)
    ibm1[0, 3] = -1.0
    half = math.sqrt(0.5)
    bend = Animation(name="bend", channels=[
        Channel(0, "rotation", AnimationSampler(np.array([0.0, 1.0]), np.array([[0.0, 0.0, 0.0, 1.0],
                                                                              [0.0, 0.0, half, half]]))),
        Channel(2, "weights", AnimationSampler(np.array([0.0, 1.0]), np.array([[0.0], [1.0]]))),
    ])
    nodes = [
        Node(translation=(1.0, 0.0, 0.0)),          # elbow
        Node(children=[0]),                         # root
        Node(mesh=0, skin=0),
        Node(light=0),
        Node(camera=0, translation=(1.0, 0.5, 4.0)),
    ]
    return Scene3D(nodes=nodes, roots=[1, 2, 3, 4], meshes=[Mesh([prim], weights=[0.0])],
                   materials=[Material(base_color=(0.2, 0.6, 0.9, 1.0), metallic=0.0, roughness=0.6)],
                   lights=[Light("directional", intensity=3.0)], cameras=[Camera(yfov=0.9)],
                   skins=[Skin(joints=[1, 0], inverse_bind_matrices=[np.eye(4), ibm1], skeleton=1)],
                   animations=[bend])
