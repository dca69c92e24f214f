"""Pins of the oracle's environment map (EnvironmentMap,
src/pathtrace.rs:1557-1741, and its use in Integrator::sample) by the
reference's own tests: the pdf integrates to 1 and equals the sampler's
pdf, samples find a bright sun (src/pathtrace.rs:3298-3345), and a white
Lambert plane under a uniform sky of radiance 1 is 0.5 under every
strategy (:3629-3667)."""
import math

import numpy as np

import oracle as O
from oxr import scenes


def sun_env():
    w, h = 32, 16
    img = np.zeros((h, w, 4), np.float32)
    img[..., :3] = 0.1
    img[..., 3] = 1.0
    img[3:5, 20:23, :3] = (50.0, 40.0, 30.0)
    return img


def env_scene(img, intensity=1.0):
    s = scenes.lit_plane(1.0, with_point=False)
    os_ = O.OracleScene(s)
    os_.set_environment(img, intensity)
    return os_


def test_environment_pdf_is_normalised_and_matches_samples():
    os_ = env_scene(sun_env())
    n = 400
    th = (np.arange(n) + 0.5) / n * math.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pdf, _ = os_.env_eval(d)
    total = float((pdf.astype(np.float64) * np.sin(T).ravel()).sum() * (math.pi / n) ** 2)
    assert abs(total - 1.0) < 0.01, total
    u = O.sobol_owen_4d(np.arange(1000, dtype=np.uint32), np.full(1000, 9, np.uint32))[:, :2]
    dirs, spdf, ok = os_.env_sample(u)
    assert ok.all()
    back, le = os_.env_eval(dirs)
    assert np.all(np.abs(spdf - back) <= 1e-3 * np.maximum(spdf, 1.0))
    assert (le[:, 0] > 10.0).sum() > 800


def test_cdf_tables_are_monotone_and_end_at_one():
    os_ = env_scene(sun_env())
    mg, cd, ms, uni = os_.env_tables()
    assert not uni and mg[0] == 0.0 and mg[-1] == 1.0 and np.all(np.diff(mg) >= 0)
    rows = cd.reshape(16, 33)
    assert np.all(rows[:, 0] == 0.0) and np.allclose(rows[:, -1], 1.0, atol=1e-6)
    assert abs(ms.astype(np.float64).sum() - 1.0) < 1e-5


def test_black_image_is_uniform_and_radiates_nothing():
    img = np.zeros((8, 16, 4), np.float32)
    os_ = env_scene(img)
    _, _, _, uni = os_.env_tables()
    assert uni
    _, le = os_.env_eval(np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0]], np.float32))
    assert np.all(le == 0.0)


def _plane_mean(os_, strategy, ambient=0.0):
    W = 6
    o = O.options(W, W, max_bounces=2, strategy=strategy, ambient=ambient)
    a = os_.render(W, W, 0, 64, o)
    return (a[..., :3].sum((0, 1)) / a[..., 3].sum())[0]


def test_uniform_environment_lights_a_lambert_plane():
    s = scenes.Scene("plane")
    m = s.add_material(scenes.lambert((0.5, 0.5, 0.5)))
    fl, _ = scenes.quad([[-1.0, 0.0, 1.0], [1.0, 0.0, 1.0], [1.0, 0.0, -1.0], [-1.0, 0.0, -1.0]])
    s.add_tris(fl, m)
    s.camera = scenes.CameraSpec((0.0, 2.0, 0.01), (0.0, 0.0, 0.0), yfov_deg=math.degrees(0.4))
    os_ = O.OracleScene(s)
    img = np.ones((8, 16, 4), np.float32)
    os_.set_environment(img, 1.0)
    for strategy in (0, 1, 2):
        assert abs(_plane_mean(os_, strategy) - 0.5) < 0.02, strategy
    os_.set_environment(None)
    assert abs(_plane_mean(os_, 0, ambient=1.0) - 0.5) < 0.02
