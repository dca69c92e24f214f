"""Pins the oracle's trace layer (a7-a13): BVH traversal against brute
force, the reference's closed forms (src/trace.rs:1222-1308) and
watertightness of shared edges."""
import ctypes as C

import numpy as np
import pytest

from oxr import _abi as A
from oxr import scenes


def _rays(o, d, t_min=0.0, t_max=np.inf):
    import oracle as O
    return O.RayBatch(o, d, t_min, t_max)


def test_box_closed_forms(oracle):
    # src/trace.rs:1222-1241 on the reference's [-1,1]^3 box.
    os_ = oracle.OracleScene(scenes.reference_box())
    rb = _rays([[0.0, 0.5, 0.9]], [[0.0, 0.0, -1.0]])
    h = os_.trace_closest(rb.arr)[0]
    t = np.frombuffer(np.uint32(h[0]).tobytes(), np.float32)[0]
    assert abs(t - 1.9) < 1e-4
    assert h[1] != A.OXR_HIT_MISS and not (h[1] & A.OXR_HIT_BACK)
    g = int(h[1] & 0x7FFFFFFF)
    u, v = np.frombuffer(np.uint32(h[2:4]).tobytes(), np.float32)
    pos, ng, _ = os_.surface(g, [1 - u - v, u, v])
    assert abs(pos[2] + 1.0) < 1e-5
    assert abs(ng[2] - 1.0) < 1e-5
    assert os_.trace_any(_rays([[0.0, 0.5, 0.9]], [[0.0, 0.0, -1.0]], 0.0, 2.0).arr)[0] == 1
    assert os_.trace_any(_rays([[0.0, 0.5, 0.9]], [[0.0, 0.0, -1.0]], 0.0, 1.5).arr)[0] == 0


def test_material_reads_factors(oracle):
    # src/trace.rs:1243-1253: the ray toward -X hits the red wall.
    s = scenes.reference_box()
    os_ = oracle.OracleScene(s)
    h = os_.trace_closest(_rays([[0.0, 0.5, 0.0]], [[-1.0, 0.0, 0.0]]).arr)[0]
    g = int(h[1] & 0x7FFFFFFF)
    m = s.materials[int(s.material_ids[g])]
    assert m.base_color[0] > 0.6 and m.base_color[1] < 0.1


def _random_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = scene.bounds()
    c = (lo + hi) / 2
    ext = (hi - lo) / 2 * 0.99
    o = c + (rng.random((n, 3)) * 2 - 1) * ext
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # Axis-parallel and signed-zero directions.
    d[: n // 16, :] = 0.0
    d[: n // 16, 0] = np.where(np.arange(n // 16) % 2 == 0, 1.0, -1.0)
    d[n // 16: n // 8, 1] = -0.0
    return o.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("make", [scenes.cornell, scenes.reference_box,
                                  lambda: scenes.material_sweep(segments=16, rings=8, floor_half=3),
                                  lambda: scenes.triangle_soup(7, 3000),
                                  # Large enough that worker threads build subtrees the build stitches.
                                  lambda: scenes.triangle_soup(8, 80000)],
                         ids=["cornell", "reference_box", "sweep", "soup3k", "soup80k"])
def test_bvh_traversal_equals_brute_force(oracle, make):
    s = make()
    os_ = oracle.OracleScene(s)
    o, d = _random_rays(s, 4000 if s.num_tris < 10000 else 1500, 11)
    rb = _rays(o, d)
    a = os_.trace_closest(rb.arr)
    b = os_.trace_closest(rb.arr, brute=True)
    assert np.array_equal(a, b)
    rb2 = _rays(o, d, 0.0, 0.7)
    assert np.array_equal(os_.trace_any(rb2.arr), os_.trace_any(rb2.arr, brute=True))
    # Closest hit with culling of single-sided back faces.
    assert np.array_equal(os_.trace_closest(rb.arr, flags=1), os_.trace_closest(rb.arr, flags=1, brute=True))


def test_closest_hit_is_nearest_and_bary_reconstructs_point(oracle):
    s = scenes.reference_box()
    os_ = oracle.OracleScene(s)
    o, d = _random_rays(s, 500, 5)
    h = os_.trace_closest(_rays(o, d).arr)
    for i in range(len(h)):
        if h[i, 1] == A.OXR_HIT_MISS:
            continue
        t = np.uint32(h[i, 0]).view(np.float32)
        g = int(h[i, 1] & 0x7FFFFFFF)
        u = np.uint32(h[i, 2]).view(np.float32)
        v = np.uint32(h[i, 3]).view(np.float32)
        pos, ng, _ = os_.surface(g, [1 - u - v, u, v])
        assert np.allclose(pos, o[i] + t * d[i], atol=2e-5)
        assert u >= -1e-6 and v >= -1e-6 and u + v <= 1 + 1e-6
        # front face <=> the ray runs against the CCW-front normal
        front = not (h[i, 1] & A.OXR_HIT_BACK)
        assert front == (float(np.dot(ng, d[i])) < 0)


def test_watertight_shared_edges_and_vertices(oracle):
    # Rays from inside a closed cube toward every vertex and edge midpoint
    # (exactly on shared edges) must hit: no leaks (Woop et al. 2013).
    s = scenes.Scene("cube")
    m = s.add_material(A.material())
    t, _ = scenes.cuboid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    s.add_tris(t, m)
    os_ = oracle.OracleScene(s)
    targets = []
    for x in (-1, 0, 1):
        for y in (-1, 0, 1):
            for z in (-1, 0, 1):
                if (x, y, z) != (0, 0, 0):
                    targets.append((x, y, z))
    origins = [(0.0, 0.0, 0.0), (0.1, -0.2, 0.3), (-0.37, 0.11, 0.05)]
    o, d = [], []
    for org in origins:
        for tg in targets:
            dd = np.array(tg, np.float64) - np.array(org)
            d.append(dd / np.linalg.norm(dd))
            o.append(org)
    h = os_.trace_closest(_rays(np.array(o), np.array(d)).arr)
    assert (h[:, 1] != A.OXR_HIT_MISS).all()


def test_offset_moves_off_the_plane_both_scales(oracle):
    # src/trace.rs:1301-1308.
    for p in ([0.0, 0.0, 0.0], [100.0, -100.0, 5.0]):
        q = oracle.offset_ray_origin(p, [0.0, 1.0, 0.0])
        assert q[1] > p[1]
        assert q[0] == p[0]
    # Far from the origin the offset is 256 ULPs along n (bit-level).
    q = oracle.offset_ray_origin([100.0, -100.0, 5.0], [0.0, 1.0, 0.0])
    assert np.float32(-100.0).view(np.int32) - np.float32(q[1]).view(np.int32) == 256


def test_barycentric_of(oracle):
    # src/trace.rs:1266-1273.
    b = oracle.barycentric_of([0.25, 0.25, 0.0], [0, 0, 0], [1, 0, 0], [0, 1, 0])
    assert abs(b[1] - 0.25) < 1e-6 and abs(b[2] - 0.25) < 1e-6
    assert oracle.barycentric_of([0.1, 0.1, 0.0], [0, 0, 0], [1, 0, 0], [2, 0, 0]) is None


def test_emissive_table(oracle):
    s = scenes.cornell()
    os_ = oracle.OracleScene(s)
    tris, pmf, cdf, lookup = os_.emissive_table()
    assert len(tris) == 2
    assert cdf[-1] == 1.0
    assert abs(pmf.sum() - 1.0) < 1e-6
    for i, g in enumerate(tris):
        assert lookup[g] == i
    assert (lookup[np.setdiff1d(np.arange(s.num_tris), tris)] == 0xFFFFFFFF).all()



def coplanar_inside(xy, p, margin=1e-6):
    """Per triangle (corners xy, (n, 3, 2), float64): 1 if point p is inside
    by more than `margin` (edge-signed distance), 0 if outside by more than
    it, -1 if within the margin of an edge."""
    a, b, c = xy[:, 0], xy[:, 1], xy[:, 2]

    def side(u, v):
        e = v - u
        ln = np.hypot(e[:, 0], e[:, 1])
        return (e[:, 0] * (p[1] - u[:, 1]) - e[:, 1] * (p[0] - u[:, 0])) / np.maximum(ln, 1e-300)
    s = np.stack([side(a, b), side(b, c), side(c, a)], 1)
    area = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    s = s * np.sign(area)[:, None]
    out = np.zeros(len(xy), np.int64)
    out[(s > margin).all(1) & (area != 0)] = 1
    out[(np.abs(s) <= margin).any(1) & (s > -margin).all(1)] = -1
    return out


def coplanar_expected(xy, owner, p):
    """The id closest hit must return for the ray down -z through p when
    the answer is fixed by the tie-break alone: p lies inside the copies of
    exactly one base triangle (identical corners, so identical t) and in no
    other triangle; the lowest of those copies' ids. None otherwise; -1
    where p is in no triangle."""
    ins = coplanar_inside(xy, p)
    if (ins < 0).any():
        return None
    hit = np.nonzero(ins == 1)[0]
    if len(hit) == 0:
        return -1
    owners = set(owner[hit].tolist())
    if len(owners) != 1 or -1 in owners:
        return None
    return int(hit.min())


def coplanar_rays(n, seed=9):
    """Rays straight down -z onto the z = 0 plane from random points."""
    rng = np.random.default_rng(seed)
    p = rng.random((n, 2))
    o = np.zeros((n, 3), np.float32)
    o[:, :2] = p
    o[:, 2] = 1.0
    d = np.zeros((n, 3), np.float32)
    d[:, 2] = -1.0
    return o, d, p


def test_coplanar_overlaps_take_the_lowest_id(oracle):
    """Equal t on coplanar overlapping triangles: closest hit returns the
    lowest global id (the oracle's stated tie-break for the unvendored
    BVH's traversal-order choice). Checked where the tie-break alone fixes
    the answer (copies with identical corners at permuted ids, closed form
    by exact 2-D containment), and the BVH equals brute force on every ray,
    including overlaps of different triangles whose t differ by rounding."""
    sc, xy, owner = scenes.coplanar_overlaps()
    os_ = oracle.OracleScene(sc)
    o, d, p = coplanar_rays(8192)
    rb = _rays(o, d)
    h_bvh = os_.trace_closest(rb.arr)
    h_bf = os_.trace_closest(rb.arr, brute=True)
    assert np.array_equal(h_bvh, h_bf)
    checked = ties = 0
    for i in range(len(p)):
        want = coplanar_expected(xy, owner, p[i])
        if want is None:
            continue
        got = -1 if h_bvh[i, 1] == A.OXR_HIT_MISS else int(h_bvh[i, 1] & 0x7FFFFFFF)
        assert got == want, (i, p[i], got, want)
        checked += 1
        ties += int(want >= 0)
    assert checked > 1000 and ties > 500, (checked, ties)
