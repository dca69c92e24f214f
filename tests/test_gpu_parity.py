"""GPU parity of the HIP path against the oracle (CPU restatement of the
reference), through the C ABI. Marked `gpu`: run on an MI355X.

Bars: Sobol' points and triangle ids bit-exact, hit t within 1 ulp;
BSDF values within the rounding of the transcendental functions;
direct-lighting images (max_bounces = 1) within 1e-3 relative per pixel;
full-bounce images within 3 sigma of the per-pixel estimator noise."""
import math

import numpy as np
import pytest

from oxr import _abi as A
from oxr import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oxr_mod():
    import oxr
    if oxr.device_count() == 0:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return oxr


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def test_sobol_owen_bit_exact(oxr_mod, oracle):
    rng = np.random.default_rng(1)
    n = 1 << 18
    idx = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    idx[:1024] = np.arange(1024)
    idx[1024:1040] = [0xFFFFFFFF, 0x80000000, 0x7FFFFFFF, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13]
    seed = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    seed[:64] = [0, 1, 77, 0xDEADBEEF] * 16
    g = oxr_mod.sobol_owen_4d(idx, seed)
    c = oracle.sobol_owen_4d(idx, seed)
    assert np.array_equal(_bits(g), _bits(c))


def _rand_rays(scene, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = scene.bounds()
    c = (lo + hi) / 2
    ext = (hi - lo) / 2 * 0.99
    o = c + (rng.random((n, 3)) * 2 - 1) * ext
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[: n // 32] = 0.0
    d[: n // 32, np.arange(n // 32) % 3] = 1.0
    d[n // 32: n // 16, 1] = -0.0
    return o.astype(np.float32), d.astype(np.float32)


def _camera_rays(scene, W, H, spp, seed=0):
    """Primary rays of samples 0..spp-1 (pattern 0), computed like the
    reference camera from the oracle's Sobol' points."""
    import oracle as O
    cam = scene.camera.resolve(W, H)
    ys, xs = np.mgrid[0:H, 0:W]
    xs = xs.reshape(-1).astype(np.uint32)
    ys = ys.reshape(-1).astype(np.uint32)
    os_, ds = [], []
    for s in range(spp):
        # pattern seed = hash_mix(pixel_seed, 0); pixel_seed = mix(mix(seed, y), x)
        def mix(a, b):
            return np.array([O.pcg_hash(int(a_) ^ O.pcg_hash(int(b_))) for a_, b_ in zip(a, b)], np.uint32)
        ps = mix(mix(np.full_like(ys, seed), ys), xs)
        pat = mix(ps, np.zeros_like(ps))
        p = O.sobol_owen_4d(np.full(len(xs), s, np.uint32), pat)
        px = xs.astype(np.float32) + p[:, 0]
        py = ys.astype(np.float32) + p[:, 1]
        ndx = (px / np.float32(W)) * np.float32(2) - np.float32(1)
        ndy = np.float32(1) - (py / np.float32(H)) * np.float32(2)
        side, up, fwd = (np.array(v, np.float32) for v in (cam.side, cam.up, cam.forward))
        dd = side[None] * (ndx * np.float32(cam.half_w))[:, None] + up[None] * (ndy * np.float32(cam.half_h))[:, None]
        dd = (dd + fwd[None]).astype(np.float32)
        dd /= np.sqrt((dd * dd).sum(1, keepdims=True)).astype(np.float32)
        os_.append(np.tile(np.array(cam.eye, np.float32), (len(xs), 1)))
        ds.append(dd)
    return np.concatenate(os_), np.concatenate(ds)


def _compare_hits(g, c):
    assert g.shape == c.shape
    assert np.array_equal(g[:, 1], c[:, 1]), "triangle ids / front faces differ"
    hit = c[:, 1] != A.OXR_HIT_MISS
    dt = np.abs(g[hit, 0].astype(np.int64) - c[hit, 0].astype(np.int64))
    assert dt.max(initial=0) <= 1, "hit t differs by more than 1 ulp"
    du = np.abs(g[hit, 2:4].astype(np.int64) - c[hit, 2:4].astype(np.int64))
    assert du.max(initial=0) <= 1


@pytest.mark.parametrize("make,n", [(scenes.cornell, 1 << 17), (scenes.reference_box, 1 << 16),
                                    (lambda: scenes.material_sweep(segments=32, rings=16, floor_half=5), 1 << 17),
                                    (lambda: scenes.triangle_soup(3, 20000), 1 << 16),
                                    (lambda: scenes.terrain(grid=96, n_emit=300, n_spheres=3000), 1 << 16)],
                         ids=["cornell", "reference_box", "sweep_small", "soup20k", "terrain_small"])
def test_closest_and_any_hit_match_oracle(oxr_mod, oracle, make, n):
    s = make()
    pt = oxr_mod.PathTracer(0)
    pt.upload(s)
    os_ = oracle.OracleScene(s)
    o, d = _rand_rays(s, n, 7)
    rb = oracle.RayBatch(o, d)
    _compare_hits(pt.trace_closest(rb.arr, n), os_.trace_closest(rb.arr))
    _compare_hits(pt.trace_closest(rb.arr, n, flags=1), os_.trace_closest(rb.arr, flags=1))
    rb2 = oracle.RayBatch(o, d, 0.0, 0.5)
    assert np.array_equal(pt.trace_any(rb2.arr, n), os_.trace_any(rb2.arr))
    pt.close()


def test_camera_ray_batches_match_oracle(oxr_mod, oracle):
    s = scenes.cornell()
    o, d = _camera_rays(s, 64, 64, 4)
    pt = oxr_mod.PathTracer(0)
    pt.upload(s)
    os_ = oracle.OracleScene(s)
    rb = oracle.RayBatch(o, d)
    _compare_hits(pt.trace_closest(rb.arr, len(o), flags=1), os_.trace_closest(rb.arr, flags=1))
    pt.close()


def test_shared_edge_rays_do_not_leak(oxr_mod, oracle):
    s = scenes.Scene("cube")
    m = s.add_material(A.material())
    t, _ = scenes.cuboid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    s.add_tris(t, m)
    pts = [(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1) if (x, y, z) != (0, 0, 0)]
    o = np.array([(0.1, -0.2, 0.3)] * len(pts), np.float32)
    d = np.array(pts, np.float64) - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    pt = oxr_mod.PathTracer(0)
    pt.upload(s)
    rb = oracle.RayBatch(o, d)
    h = pt.trace_closest(rb.arr, len(o))
    assert (h[:, 1] != A.OXR_HIT_MISS).all()
    _compare_hits(h, oracle.OracleScene(s).trace_closest(rb.arr))
    pt.close()


def _bsdf_cases():
    ms = A.material_sample
    return [
        ms((0.8, 0.5, 0.2), 0.0, 0.8), ms((0.8, 0.8, 0.8), 0.0, 0.45), ms((0.95, 0.64, 0.54), 1.0, 0.6),
        ms((0.9, 0.9, 0.9), 0.5, 0.5), ms((1, 1, 1), 0.0, 0.5, transmission=1.0),
        ms((1, 1, 1), 0.0, 0.5, transmission=1.0, thickness=1.0),
        ms((0.2, 0.3, 0.8), 0.0, 0.7, clearcoat=1.0, clearcoat_roughness=0.4),
        ms((0.3, 0.1, 0.1), 0.0, 0.8, sheen_color=(0.9, 0.7, 0.7), sheen_roughness=0.6),
        ms((0.95, 0.9, 0.8), 1.0, 0.5, anisotropy=0.9, anisotropy_direction=(1, 0, 0)),
        ms((0.8, 0.8, 0.8), 0.0, 0.3, iridescence=1.0, iridescence_thickness=550.0),
        ms((0.3, 0.6, 0.2), 0.0, 0.5, diffuse_transmission=0.6, diffuse_transmission_color=(0.9, 0.5, 0.3),
           transmission=0.3),
        ms((1, 1, 1), 0.0, 1.0, specular=0.0),
    ]


@pytest.mark.parametrize("ci", range(12))
def test_bsdf_eval_and_sample_match_oracle(oxr_mod, oracle, ci):
    ms = _bsdf_cases()[ci]
    rng = np.random.default_rng(ci)
    n = 20000
    v = rng.normal(size=(n, 3)).astype(np.float32)
    v[:, 2] = np.abs(v[:, 2]) + 0.05
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    l = rng.normal(size=(n, 3)).astype(np.float32)
    l /= np.linalg.norm(l, axis=1, keepdims=True)
    ng = np.tile(np.array([0, 0, 1], np.float32), (n, 1))
    front = (rng.random(n) < 0.8).astype(np.uint8)
    fg, pg = oxr_mod.bsdf_eval(ms, v, ng, front, l)
    fc, pc = oracle.bsdf_eval(ms, v, ng, front, l)
    tol = lambda x: 2e-5 * np.maximum(np.abs(x), 1e-3)
    assert np.all(np.abs(fg - fc) <= tol(fc)), np.abs(fg - fc).max()
    assert np.all(np.abs(pg - pc) <= tol(pc))
    u = oracle.sobol_owen_4d(np.arange(n, dtype=np.uint32), np.full(n, 99, np.uint32))[:, :3]
    lg, fg, pg, okg = oxr_mod.bsdf_sample(ms, v, ng, front, u)
    lc, fc, pc, okc = oracle.bsdf_sample(ms, v, ng, front, u)
    agree = okg == okc
    assert agree.mean() > 0.999
    both = okg & okc
    assert np.abs(lg[both] - lc[both]).max(initial=0) < 1e-4
    close = np.all(np.abs(fg[both] - fc[both]) <= 1e-3 * np.maximum(np.abs(fc[both]), 1e-3), axis=1)
    assert close.mean() > 0.999


def _render_gpu(oxr_mod, s, W, spp, bounces, **kw):
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=W, height=W, ambient=getattr(s, "ambient", 0.0),
                                 background=(0, 0, 0, 255),
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=spp, max_bounces=bounces, **kw))
    pt.sync(s, opts)
    pt.refine(spp)
    acc, sq = pt.accum(with_sum_sq=True)
    return pt, acc, sq


def _render_cpu(oracle, s, W, spp, bounces, **kw):
    os_ = oracle.OracleScene(s)
    o = oracle.options(W, W, spp=spp, max_bounces=bounces, ambient=getattr(s, "ambient", 0.0), **kw)
    return os_.render(W, W, 0, spp, o, sum_sq=True)


@pytest.mark.parametrize("make,W,spp", [(scenes.cornell, 64, 16), (scenes.reference_box, 48, 16),
                                        (lambda: scenes.material_sweep(segments=32, rings=16, floor_half=6), 48, 8),
                                        (scenes.glossy_caustic, 48, 8),
                                        (lambda: scenes.terrain(grid=96, n_emit=300, n_spheres=3000), 48, 8)],
                         ids=["cornell", "reference_box", "sweep_small", "caustic", "terrain_small"])
def test_direct_lighting_matches_oracle_per_pixel(oxr_mod, oracle, make, W, spp):
    s = make()
    _, ag, _ = _render_gpu(oxr_mod, s, W, spp, 1)
    ac, _ = _render_cpu(oracle, s, W, spp, 1)
    assert np.array_equal(ag[..., 3], ac[..., 3]), "coverage differs"
    cov = ac[..., 3] > 0
    mg = ag[..., :3][cov] / ag[..., 3][cov][:, None]
    mc = ac[..., :3][cov] / ac[..., 3][cov][:, None]
    rel = np.abs(mg - mc) / np.maximum(np.abs(mc), 1e-3)
    print(f"direct-lighting rel err: max {rel.max():.3g}, pixels > 1e-3: {(rel > 1e-3).mean():.2e}")
    assert rel.max() <= 1e-3, rel.max()


def _many_lights(n_lights):
    """A Lambert floor and box lit by n_lights coloured point and spot
    lights and a small emitter: n_lights + 1 NEE slots per path, so the
    shading kernel stages its shadow rays in LDS up to 4 slots and appends
    them per light beyond."""
    from oxr import _abi as A
    s = scenes.Scene(f"lights{n_lights}")
    fl = s.add_material(scenes.lambert((0.7, 0.7, 0.7)))
    t, nr = scenes.quad([[-3, 0, 3], [3, 0, 3], [3, 0, -3], [-3, 0, -3]])
    s.add_tris(t, fl, normals=nr)
    bx = s.add_material(A.material((0.6, 0.5, 0.4, 1.0), metallic=0.2, roughness=0.4))
    t, nr = scenes.cuboid((-0.5, 0.0, -0.5), (0.5, 1.0, 0.5))
    s.add_tris(t, bx, normals=nr)
    em = s.add_material(scenes.lambert((0.0, 0.0, 0.0), emissive=(6.0, 6.0, 5.0)))
    t, nr = scenes.quad([[-0.3, 2.5, 0.3], [0.3, 2.5, 0.3], [0.3, 2.5, -0.3], [-0.3, 2.5, -0.3]])
    s.add_tris(t[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]], em)  # facing down
    for k in range(n_lights):
        a = 2.0 * math.pi * k / n_lights
        pos = (2.0 * math.cos(a), 2.0 + 0.3 * k, 2.0 * math.sin(a))
        col = (1.0, 0.5 + 0.1 * k, 1.0 - 0.15 * k)
        if k % 2:
            d = [-pos[0], -pos[1], -pos[2]]
            nd = math.sqrt(sum(v * v for v in d))
            s.lights.append(A.light(A.OXR_LIGHT_SPOT, color=col, intensity=6.0, position=pos,
                                    direction=tuple(v / nd for v in d), inner=0.3, outer=0.7))
        else:
            s.lights.append(A.light(A.OXR_LIGHT_POINT, color=col, intensity=5.0, position=pos))
    s.camera = scenes.CameraSpec((0.0, 3.0, 4.5), (0.0, 0.4, 0.0), yfov_deg=50.0)
    s.ambient = 0.2
    return s


@pytest.mark.parametrize("n_lights", [3, 5])
def test_many_light_slots_match_oracle(oxr_mod, oracle, n_lights):
    """Direct lighting per pixel and full bounces within 3 sigma with 4
    (staged) and 6 (appended per light) NEE slots per path."""
    s = _many_lights(n_lights)
    W = 40
    _, ag, _ = _render_gpu(oxr_mod, s, W, 8, 1)
    ac, _ = _render_cpu(oracle, s, W, 8, 1)
    assert np.array_equal(ag[..., 3], ac[..., 3])
    cov = ac[..., 3] > 0
    mg = ag[..., :3][cov] / ag[..., 3][cov][:, None]
    mc = ac[..., :3][cov] / ac[..., 3][cov][:, None]
    rel = np.abs(mg - mc) / np.maximum(np.abs(mc), 1e-3)
    assert rel.max() <= 1e-3, rel.max()
    _, ag, sqg = _render_gpu(oxr_mod, s, 24, 32, 6, seed=3)
    ac, sqc = _render_cpu(oracle, s, 24, 32, 6, seed=3)
    cov = ac[..., 3] > 0
    n = ac[..., 3][cov][:, None]
    mg, mc = ag[..., :3][cov] / n, ac[..., :3][cov] / n
    sigma = np.sqrt(np.maximum(sqc[cov] / n - mc ** 2, 0.0) / n) + 1e-6
    assert (np.abs(mg - mc) / sigma > 3).mean() < 0.01


@pytest.mark.parametrize("make,W,spp,b", [(scenes.cornell, 32, 64, 8), (scenes.reference_box, 32, 64, 6),
                                          (scenes.glossy_caustic, 24, 32, 8),
                                          (lambda: scenes.material_sweep(segments=32, rings=16, floor_half=6), 32, 32,
                                           8)],
                         ids=["cornell", "reference_box", "caustic", "sweep_small"])
def test_full_bounce_images_agree_within_3_sigma(oxr_mod, oracle, make, W, spp, b):
    s = make()
    _, ag, sqg = _render_gpu(oxr_mod, s, W, spp, b, seed=5)
    ac, sqc = _render_cpu(oracle, s, W, spp, b, seed=5)
    cov = ac[..., 3] > 0
    n = ac[..., 3][cov][:, None]
    mg = ag[..., :3][cov] / n
    mc = ac[..., :3][cov] / n
    var = np.maximum(sqc[cov] / n - mc ** 2, 0.0) / n
    sigma = np.sqrt(var) + 1e-6
    z = np.abs(mg - mc) / sigma
    print(f"max |gpu-cpu|/sigma = {z.max():.3g}; frac > 3 sigma = {(z > 3).mean():.3g}")
    # Same seeds, same sample sequences: most pixels agree far below 3 sigma.
    assert (z > 3).mean() < 0.01


@pytest.mark.parametrize("make", [scenes.cornell, scenes.reference_box,
                                  lambda: scenes.material_sweep(segments=32, rings=16, floor_half=6)],
                         ids=["cornell", "reference_box", "sweep_small"])
def test_shading_without_extension_lobes_is_bit_identical(oxr_mod, make):
    """Scenes whose materials have no clearcoat, sheen, iridescence,
    transmission, diffuse transmission or anisotropy shade with the variant
    that leaves those lobes out; one unused extension material in the table
    selects the full variant. Both accumulate the same bits."""
    s = make()
    _, a, sa = _render_gpu(oxr_mod, s, 40, 8, 8)
    full = make()
    full.add_material(A.material((1.0, 1.0, 1.0, 1.0), clearcoat=1.0, sheen_color=(0.5, 0.5, 0.5), transmission=0.5,
                                 diffuse_transmission=0.3, iridescence=1.0))
    assert full.num_tris == s.num_tris
    _, b, sb = _render_gpu(oxr_mod, full, 40, 8, 8)
    assert np.array_equal(a, b)
    assert np.array_equal(sa, sb)


@pytest.mark.parametrize("make", [scenes.glossy_caustic, lambda: scenes.terrain(grid=48, n_emit=200, n_spheres=300)],
                         ids=["caustic", "terrain_small"])
def test_shading_without_iridescence_is_bit_identical(oxr_mod, make):
    """Scenes with extension lobes (transmission; clearcoat, sheen) but no
    iridescent material shade with the variant that leaves the thin-film
    Fresnel out; one unused iridescent material in the table selects the
    variant with every lobe. Both accumulate the same bits."""
    s = make()
    _, a, sa = _render_gpu(oxr_mod, s, 40, 8, 8)
    full = make()
    full.add_material(A.material((1.0, 1.0, 1.0, 1.0), iridescence=1.0))
    assert full.num_tris == s.num_tris
    _, b, sb = _render_gpu(oxr_mod, full, 40, 8, 8)
    assert np.array_equal(a, b)
    assert np.array_equal(sa, sb)


@pytest.mark.parametrize("bounces,camera", [(0, "perspective"), (3, "perspective"), (0, "axis_ortho"),
                                            (3, "axis_ortho")])
def test_traversal_counts_match_the_oracle_restatement(oxr_mod, oracle, bounces, camera):
    """The counting closest-hit variant (the roofline's node and triangle
    fetches) against the oracle's restatement of the device traversal
    (oxrcpu_bvh4_count) over the exported BVH and the oracle's own path rays
    of the same samples: ray counts equal; fetches within 1% (ties in
    child distance may order children differently). axis_ortho: an
    orthographic camera looking down -z, so every camera ray has two zero
    direction components (and mirror bounces off axis-aligned faces keep
    them): the slab reciprocal of such components (+-2^64 on both sides)
    decides whether those axes cull; an infinite one fetches several times
    the nodes."""
    s = scenes.material_sweep(segments=32, rings=16, floor_half=6)
    if camera == "axis_ortho":
        s.camera = scenes.CameraSpec((0.0, 1.0, 9.0), (0.0, 1.0, 0.0), projection=A.OXR_PROJ_ORTHOGRAPHIC, ymag=3.0)
    W, spp = 48, 4
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=W, height=W, ambient=s.ambient,
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=spp, max_bounces=bounces))
    pt.sync(s, opts)
    pt.set_timing(False, count=True)
    pt.refine(spp)
    st = pt.stats()
    nodes, tris = pt.bvh_export()
    pt.close()
    rays = oracle.OracleScene(s).record_rays(W, W, 0, spp, 0, W, max_rays=1 << 21,
                                             opts=oracle.options(W, W, max_bounces=bounces, ambient=s.ambient))
    c = oracle.bvh4_count(nodes, tris, rays).astype(np.int64)
    if bounces == 0:
        assert st["closest_rays"] == len(rays) == W * W * spp
    else:
        assert abs(st["closest_rays"] - len(rays)) <= 0.001 * len(rays)
    assert st["closest_nodes"] > 0 and st["closest_tris"] > 0
    assert abs(st["closest_nodes"] - c[:, 0].sum()) <= 0.01 * c[:, 0].sum()
    assert abs(st["closest_tris"] - c[:, 1].sum()) <= 0.01 * c[:, 1].sum()


def test_occlusion_counts_match_the_oracle_restatement(oxr_mod, oracle):
    """The counting any-hit variant (the occlusion roofline's node and
    triangle fetches) against the oracle's restatement of the device's
    occlusion order (oxrcpu_bvh4_count_any: leaves first, inner children
    near to far, stop at the first triangle hit) over the oracle's own
    occlusion rays of the same samples: ray counts and fetches within 1%."""
    s = scenes.material_sweep(segments=32, rings=16, floor_half=6)
    W, spp, bounces = 48, 4, 3
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=W, height=W, ambient=s.ambient,
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=spp, max_bounces=bounces))
    pt.sync(s, opts)
    pt.set_timing(False, count=True)
    pt.refine(spp)
    st = pt.stats()
    nodes, tris = pt.bvh_export()
    pt.close()
    rays = oracle.OracleScene(s).record_shadow_rays(W, W, 0, spp, 0, W, max_rays=1 << 21,
                                                    opts=oracle.options(W, W, max_bounces=bounces, ambient=s.ambient))
    c = oracle.bvh4_count_any(nodes, tris, rays).astype(np.int64)
    assert abs(st["shadow_rays"] - len(rays)) <= 0.001 * len(rays)
    assert st["anyhit_nodes"] > 0 and st["anyhit_tris"] > 0
    assert abs(st["anyhit_nodes"] - c[:, 0].sum()) <= 0.01 * c[:, 0].sum()
    assert abs(st["anyhit_tris"] - c[:, 1].sum()) <= 0.01 * c[:, 1].sum()


def test_refinement_is_additive_and_deterministic(oxr_mod):
    s = scenes.area_light_floor(0.5, 0.0)
    pt, a, _ = _render_gpu(oxr_mod, s, 9, 8, 4)
    pt2, b, _ = _render_gpu(oxr_mod, s, 9, 8, 4)
    assert np.array_equal(a, b)
    pt3 = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=9, height=9, ambient=0.0,
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=8, max_bounces=4))
    pt3.sync(s, opts)
    pt3.refine(3)
    pt3.refine(5)
    assert pt3.samples() == 8
    assert np.array_equal(pt3.accum(), a)
    assert not pt3.sync(s, opts)
    seeded = oxr_mod.RenderOptions(width=9, height=9, ambient=0.0,
                                   path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=8, max_bounces=4, seed=1))
    assert pt3.sync(s, seeded)
    assert pt3.samples() == 0


def test_gpu_accumulator_equals_oracle_sample_sums(oxr_mod, oracle):
    # Same samples: the f64 sums agree to float32 rounding of each sample.
    s = scenes.cornell()
    _, ag, _ = _render_gpu(oxr_mod, s, 32, 8, 8)
    ac, _ = _render_cpu(oracle, s, 32, 8, 8)
    assert np.array_equal(ag[..., 3], ac[..., 3])
    close = np.abs(ag[..., :3] - ac[..., :3]) <= 1e-4 * np.maximum(np.abs(ac[..., :3]), 1.0)
    print("pixels with equal sums:", close.mean())
    assert close.mean() > 0.95


def test_resolve_matches_oracle_resolve(oxr_mod, oracle):
    s = scenes.area_light_floor(0.5, 0.0)
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=16, height=16, ambient=0.0, background=(200, 10, 77, 128),
                                 tone_map=oxr_mod.ACES, exposure=1.5,
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=4, max_bounces=2))
    s.camera = scenes.CameraSpec((0.0, 1.0, 6.0), (0.0, 0.0, 0.0), yfov_deg=50.0)
    pt.sync(s, opts)
    pt.refine(4)
    acc = pt.accum()
    img = pt.image()
    ref = oracle.resolve_rgba8(acc, 4, background=(200, 10, 77, 128), tone_map=oxr_mod.ACES, exposure=1.5)
    assert np.array_equal(img, ref)
    assert list(img[0, 0]) == [200, 10, 77, 128]
    hdr = pt.hdr()
    assert np.array_equal(hdr, oracle.resolve_hdr(acc, 4, background=(200, 10, 77, 128)))


def test_resolve_alpha_rounds_half_away_from_zero(oxr_mod, oracle):
    """PathTracer::image (src/pathtrace.rs:2694-2716): alpha bytes use
    f32::round. At 6 spp over a silhouette on a transparent background a
    pixel with one covered sample has alpha 255/6 = 42.5, which rounds to
    43; the device bytes equal the oracle's, which the resolve pins fix."""
    s = scenes.Scene("disc")
    m = s.add_material(scenes.lambert((0.8, 0.8, 0.8), emissive=(1.0, 0.5, 0.25)))
    t, _ = scenes.uv_sphere(1.0, 24, 12)
    s.add_tris(t, m)
    s.camera = scenes.CameraSpec((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), yfov_deg=40.0)
    W, spp = 48, 6
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=W, height=W, ambient=0.0, background=(0, 0, 0, 0),
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=spp, max_bounces=1))
    pt.sync(s, opts)
    pt.refine(spp)
    acc = pt.accum()
    img = pt.image()
    pt.close()
    ref = oracle.resolve_rgba8(acc, spp, background=(0, 0, 0, 0))
    assert np.array_equal(img, ref)
    one = acc[..., 3] == 1
    assert one.any(), "no pixel with exactly one covered sample"
    assert (img[..., 3][one] == 43).all()
    assert np.array_equal(pt_hdr_alpha(acc, spp), img[..., 3])


def pt_hdr_alpha(acc, spp):
    """Alpha bytes of the coverage fraction, f32 and half away from zero."""
    f = np.clip((acc[..., 3] / spp).astype(np.float32), 0, 1) * np.float32(255.0)
    return np.floor(f.astype(np.float64) + 0.5).astype(np.uint8)


# ---- the reference's end-to-end closed forms, on the GPU --------------------

def _mean(oxr_mod, s, W, spp, b, **kw):
    amb = kw.pop("ambient", 0.0)
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=W, height=W, ambient=amb, background=(0, 0, 0, 255),
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=spp, max_bounces=b, **kw))
    pt.sync(s, opts)
    pt.refine(spp)
    return pt.hdr()[..., :3].reshape(-1, 3).astype(np.float64).mean(0)


def test_gpu_white_furnace(oxr_mod):
    for st in (0, 1, 2):
        m = _mean(oxr_mod, scenes.furnace(0.5, 1.0), 8, 64, 64, strategy=st)
        assert np.all(np.abs(m - 2.0) < 0.04), (st, m)


def test_gpu_point_light_and_clamp(oxr_mod):
    s = scenes.lit_plane(1.0)
    assert abs(_mean(oxr_mod, s, 2, 4, 1)[0] - 4.0 / math.pi) < 0.01
    assert abs(_mean(oxr_mod, s, 2, 4, 1, clamp=0.5)[0] - 0.5) < 1e-3


def test_gpu_smooth_glass(oxr_mod):
    assert abs(_mean(oxr_mod, scenes.glass_cube(False), 4, 16, 8)[0] - 0.92) < 0.03
    assert abs(_mean(oxr_mod, scenes.glass_cube(True), 4, 16, 8)[0] - 0.46) < 0.03


def test_gpu_mis_light_bsdf_agree(oxr_mod):
    for rough, metal in ((1.0, 0.0), (0.35, 0.5)):
        s = scenes.area_light_floor(rough, metal)
        means = [_mean(oxr_mod, s, 12, 256, 1, strategy=st)[1] for st in (0, 1, 2)]
        assert all(abs(m - means[0]) < 0.03 * means[0] for m in means), means


def test_gpu_empty_scene_is_background(oxr_mod):
    s = scenes.Scene("empty")
    s.camera = scenes.CameraSpec((0, 0, 3), (0, 0, 0))
    pt = oxr_mod.PathTracer(0)
    opts = oxr_mod.RenderOptions(width=8, height=8, background=(200, 10, 77, 128),
                                 path_trace=oxr_mod.PathTraceOptions(samples_per_pixel=2, max_bounces=2))
    pt.sync(s, opts)
    pt.refine(2)
    assert (pt.image().reshape(-1, 4) == [200, 10, 77, 128]).all()


def test_render_window_equals_the_frame_pixels(oxr_mod):
    """oxr_pt_set_window: refine over a window traces those pixels with the
    samples they have in the whole frame, so the window's accumulator (and
    sums of squares) equal the frame's over the same pixels bit for bit."""
    s = scenes.material_sweep(segments=32, rings=16, floor_half=6)
    W, spp = 96, 8
    pt, acc, sq = _render_gpu(oxr_mod, s, W, spp, 8)
    for x0, y0, w, h in ((17, 30, 40, 23), (0, 0, 96, 5), (90, 91, 6, 5)):
        pt.set_window(x0, y0, w, h)
        pt.refine(3)
        pt.refine(spp - 3)
        a, q = pt.accum(with_sum_sq=True)
        assert a.shape == (h, w, 4)
        assert np.array_equal(a, acc[y0:y0 + h, x0:x0 + w]), (x0, y0, w, h)
        assert np.array_equal(q, sq[y0:y0 + h, x0:x0 + w])
    pt.set_window()
    pt.refine(spp)
    assert np.array_equal(pt.accum(), acc)
    with pytest.raises(oxr_mod.InvalidOptions):  # outside the frame: an invalid argument
        pt.set_window(90, 90, 10, 10)
    pt.refine(1)  # the rejected window left the whole frame in place
    assert pt.accum().shape == (W, W, 4)
    pt.close()
