"""A prepared scene graph on the GPU against the oracle: the example graph
(strip floor, mirrored indexed box with COLOR_0, a normal-mapped panel with
generated tangents, an emissive fan, spot and point lights on transformed
nodes, seen through its camera node) goes through the prepare layer
(src/prepare.rs:1098-1237, src/camera.rs:177-332) into `sync`, and its
direct-lighting image matches the oracle's per pixel, its full-bounce
image within 3 sigma. The reference's skinned, morphed beam goes through
the animated pipeline: rays at its rest and bent tips hit as its
scanline test asserts (tests/scanline_pbr.rs:252-270), and its posed
direct lighting matches the oracle per pixel."""
import numpy as np
import pytest

import oracle as O
from oxr import prepare as P

pytestmark = pytest.mark.gpu


def _scene():
    s = P.to_scene(P.build(P.showcase_graph()), "graph", scene_camera=0)
    s.ambient = 0.1
    return s


def _pair(W, spp, bounces, s=None):
    import oxr
    s = _scene() if s is None else s
    pt = oxr.PathTracer(0)
    o = oxr.RenderOptions(width=W, height=W, ambient=s.ambient)
    o.path_trace.samples_per_pixel = spp
    o.path_trace.max_bounces = bounces
    pt.sync(s, o)
    pt.refine(spp)
    ag, sg = pt.accum(with_sum_sq=True)
    pt.close()
    ac, sc = O.OracleScene(s).render(W, W, 0, spp, O.options(W, W, max_bounces=bounces, ambient=s.ambient),
                                     sum_sq=True)
    return ag, sg, ac, sc


def test_prepared_graph_direct_lighting_matches_oracle():
    ag, _, ac, _ = _pair(64, 8, 1)
    assert np.array_equal(ag[..., 3], ac[..., 3])
    cov = ac[..., 3] > 0
    mg = ag[..., :3][cov] / ag[..., 3][cov][:, None]
    mc = ac[..., :3][cov] / ac[..., 3][cov][:, None]
    assert mc.max() > 0.1
    rel = np.abs(mg - mc) / np.maximum(np.abs(mc), 1e-3)
    print(f"direct lighting: max rel {rel.max():.3g}, > 1e-3: {(rel > 1e-3).mean():.2e}")
    assert rel.max() <= 1e-3, rel.max()


def test_prepared_graph_full_bounce_within_3_sigma():
    W, spp = 40, 64
    ag, sg, ac, sc = _pair(W, spp, 6)
    n = np.maximum(ac[..., 3:4], 1.0)
    mg, mc = ag[..., :3] / n, ac[..., :3] / n
    var = (np.maximum(sg / n - mg ** 2, 0) + np.maximum(sc / n - mc ** 2, 0)) / n
    z = np.abs(mg - mc) / np.sqrt(var + 1e-20)
    assert (z > 3.0).mean() < 0.01


def test_animated_beam_tips_and_direct_lighting():
    import oxr
    from oxr import _abi as A
    for t, hit, miss in ((0.0, (1.9, 0.0), (0.72, 0.95)), (1.0, (0.72, 0.95), (1.9, 0.0))):
        s = P.to_scene(P.build(P.skinned_morph_beam(), time=t), "beam", scene_camera=0)
        pt = oxr.PathTracer(0)
        pt.upload(s)
        o = np.array([[hit[0], hit[1], 4.0], [miss[0], miss[1], 4.0]], np.float32)
        d = np.array([[0.0, 0.0, -1.0]] * 2, np.float32)
        h = pt.trace_closest(O.RayBatch(o, d).arr, 2)
        pt.close()
        assert h[0, 1] != A.OXR_HIT_MISS and h[1, 1] == A.OXR_HIT_MISS, t
    s = P.to_scene(P.build(P.skinned_morph_beam(), time=0.5), "beam", scene_camera=0)
    s.ambient = 0.05
    ag, _, ac, _ = _pair(96, 4, 1, s)
    assert np.array_equal(ag[..., 3], ac[..., 3])
    cov = ac[..., 3] > 0
    assert cov.sum() > 200
    mg = ag[..., :3][cov] / ag[..., 3][cov][:, None]
    mc = ac[..., :3][cov] / ac[..., 3][cov][:, None]
    rel = np.abs(mg - mc) / np.maximum(np.abs(mc), 1e-3)
    assert rel.max() <= 1e-3, rel.max()


def test_sync_prepares_scene_graphs_with_their_options():
    """PathTracer.sync over a scene graph (src/pathtrace.rs:2474-2508):
    prepared as an explicit prepare call would be (same accumulators bit for
    bit), prepared again and reset exactly when a preparation input (here
    the animation time) changes, untouched when nothing changed; and the
    Renderer takes a scene graph too."""
    import oxr
    graph = P.skinned_morph_beam()
    W = 48

    def opts(**kw):
        o = oxr.RenderOptions(width=W, height=W, ambient=0.05, scene_camera=0, **kw)
        o.path_trace.samples_per_pixel = 4
        o.path_trace.max_bounces = 2
        return o
    pt = oxr.PathTracer(0)
    assert pt.sync(graph, opts(time=1.0))
    assert not pt.sync(graph, opts(time=1.0))
    pt.refine(4)
    bent = pt.accum()
    assert pt.sync(graph, opts(time=0.5))
    assert pt.samples() == 0
    pt.refine(4)
    half = pt.accum()
    pt.close()
    assert not np.array_equal(bent, half)
    ref = oxr.PathTracer(0)
    ref.sync(P.to_scene(P.build(graph, time=0.5), scene_camera=0), opts(time=0.5))
    ref.refine(4)
    assert np.array_equal(half, ref.accum())
    ref.close()
    img = oxr.PathTraceRenderer(0).render(graph, opts(time=1.0))
    assert img.shape[:2] == (W, W) and img[..., 3].max() > 0
