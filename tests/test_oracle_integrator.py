"""Pins the oracle's integrator (a23-a28) with the reference's end-to-end
closed forms (src/pathtrace.rs:3346-3750)."""
import math

import numpy as np

from oxr import scenes


def mean_rgb(oracle, s, w, spp, bounces, **kw):
    os_ = oracle.OracleScene(s)
    o = oracle.options(w, w, spp=spp, max_bounces=bounces, ambient=kw.pop("ambient", 0.0), **kw)
    acc = os_.render(w, w, 0, spp, o)
    img = oracle.resolve_hdr(acc, spp)
    return img[..., :3].reshape(-1, 3).astype(np.float64).mean(0)


def test_white_furnace_converges_to_analytic_radiance(oracle):
    s = scenes.furnace(0.5, 1.0)
    for strategy in (0, 1, 2):
        m = mean_rgb(oracle, s, 8, 64, 64, strategy=strategy)
        assert np.all(np.abs(m - 2.0) < 0.04), (strategy, m)


def test_mis_light_only_and_bsdf_only_agree(oracle):
    for rough, metal in ((1.0, 0.0), (0.35, 0.5)):
        s = scenes.area_light_floor(rough, metal)
        means = [mean_rgb(oracle, s, 12, 256, 1, strategy=st)[1] for st in (0, 1, 2)]
        ref = means[0]
        assert ref > 0.05
        for m in means:
            assert abs(m - ref) < 0.03 * ref, (rough, means)


def test_error_decreases_with_sample_count(oracle):
    s = scenes.area_light_floor(0.5, 0.0)
    os_ = oracle.OracleScene(s)
    ref = oracle.resolve_hdr(os_.render(10, 10, 0, 1024, oracle.options(10, 10, max_bounces=3)), 1024)
    errs = []
    for spp in (4, 16, 64):
        img = oracle.resolve_hdr(os_.render(10, 10, 0, spp, oracle.options(10, 10, max_bounces=3, seed=99)), spp)
        errs.append(float(np.sqrt(((img - ref).astype(np.float64) ** 2).mean())))
    assert errs[1] < errs[0] * 0.8 and errs[2] < errs[1] * 0.8, errs


def test_cornell_box_red_wall_bleeds_onto_floor(oracle):
    # src/pathtrace.rs:3465-3490 on the reference's test box.
    s = scenes.reference_box()
    os_ = oracle.OracleScene(s)

    def red_ratio(b):
        img = oracle.resolve_hdr(os_.render(32, 32, 0, 64, oracle.options(32, 32, max_bounces=b)), 64)
        avg = lambda x0: img[28:31, x0:x0 + 3, :3].reshape(-1, 3).sum(0)
        l, r = avg(4), avg(25)
        return (l[0] / l[1]) / (r[0] / r[1])

    assert abs(red_ratio(1) - 1.0) < 0.05
    assert red_ratio(6) > 1.15


def test_renders_are_deterministic_and_refinement_is_additive(oracle):
    s = scenes.area_light_floor(0.5, 0.0)
    os_ = oracle.OracleScene(s)
    o = oracle.options(9, 9, max_bounces=4)
    a = os_.render(9, 9, 0, 8, o)
    b = os_.render(9, 9, 0, 8, o)
    assert np.array_equal(a, b)
    c = os_.render(9, 9, 0, 3, o) + 0.0
    # refine(3) then refine(5) continues the sample index into the same sums.
    acc = np.zeros_like(a)
    acc += os_.render(9, 9, 0, 3, o)
    # Continue in place (the oracle adds into the caller's buffer).
    import ctypes as C
    L = oracle.lib()
    cam = s.camera.resolve(9, 9)
    L.oxrcpu_render(os_.h, C.byref(cam), C.byref(o), 3, 5, 0, 9, 0, acc.ctypes.data_as(oracle.dp), oracle.dp())
    assert np.array_equal(acc, a)
    seeded = oracle.options(9, 9, max_bounces=4, seed=1)
    assert not np.array_equal(os_.render(9, 9, 0, 8, seeded), a)
    assert c.shape == a.shape


def test_primary_misses_keep_background_bytes(oracle):
    # src/pathtrace.rs:3608-3626: an uncovered pixel resolves to the
    # background bytes exactly.
    acc = np.zeros((2, 2, 4))
    acc[0, 0] = [0.5, 0.5, 0.5, 4]
    img = oracle.resolve_rgba8(acc, 4, background=(200, 10, 77, 128))
    assert list(img[1, 1]) == [200, 10, 77, 128]
    assert list(img[0, 1]) == [200, 10, 77, 128]
    assert img[0, 0, 3] == 255


def test_uniform_sky_lights_a_lambert_plane(oracle):
    # src/pathtrace.rs:3628-3667 (ambient-sky variant): L = rho * 1.
    s = scenes.lit_plane(0.5, with_point=False)
    s.camera = scenes.CameraSpec((0.0, 2.0, 0.01), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.4))
    m = mean_rgb(oracle, s, 6, 64, 2, ambient=1.0)
    assert abs(m[0] - 0.5) < 0.02, m


def test_smooth_glass_transmits_with_fresnel_and_beer_lambert(oracle):
    thin = mean_rgb(oracle, scenes.glass_cube(False), 4, 16, 8)[0]
    vol = mean_rgb(oracle, scenes.glass_cube(True), 4, 16, 8)[0]
    assert abs(thin - 0.92) < 0.03, thin
    assert abs(vol - 0.46) < 0.03, vol


def test_point_lights_and_clamp(oracle):
    s = scenes.lit_plane(1.0)
    l = mean_rgb(oracle, s, 2, 4, 1)[0]
    assert abs(l - 4.0 / math.pi) < 0.01, l
    c = mean_rgb(oracle, s, 2, 4, 1, clamp=0.5)[0]
    assert abs(c - 0.5) < 1e-3, c
