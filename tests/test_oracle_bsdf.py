"""Pins the oracle's BSDF (a14-a19) and light sampling (a20-a22) with the
reference's quadrature / Monte Carlo / closed-form tests
(src/pathtrace.rs:2954-3296)."""
import math

import numpy as np
import pytest

from oxr import _abi as A


def _norm(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v)).astype(np.float32)


def quadrature(oracle, ms, v, ng, front, nt=360, np_=180):
    th = (np.arange(nt) + 0.5) / nt * math.pi
    ph = (np.arange(np_) + 0.5) / np_ * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    l = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    dw = (np.sin(T) * (math.pi / nt) * (math.pi / np_) * 2.0).reshape(-1)
    f, pdf = oracle.bsdf_eval(ms, v, ng, front, l)
    return (f.astype(np.float64) * dw[:, None]).sum(0), float((pdf.astype(np.float64) * dw).sum())


def monte_carlo(oracle, ms, v, ng, front, n):
    u = oracle.sobol_owen_4d(np.arange(n, dtype=np.uint32), np.full(n, 12345, np.uint32))[:, :3]
    _, f, pdf, ok = oracle.bsdf_sample(ms, v, ng, front, u)
    w = np.zeros((n, 3))
    w[ok] = (f[ok] / pdf[ok, None]).astype(np.float64)
    return w.sum(0) / n


def _cases_core():
    ms = A.material_sample
    cases = [
        ("rough dielectric", ms((0.8, 0.5, 0.2), 0.0, 0.8), True),
        ("glossy dielectric", ms((0.8, 0.8, 0.8), 0.0, 0.45), True),
        ("rough metal", ms((0.95, 0.64, 0.54), 1.0, 0.6), True),
        ("half metal", ms((0.9, 0.9, 0.9), 0.5, 0.5), True),
        ("thin transmission", ms((1, 1, 1), 0.0, 0.5, transmission=1.0), True),
        ("volume entering", ms((1, 1, 1), 0.0, 0.5, transmission=1.0, thickness=1.0), True),
        ("volume leaving", ms((1, 1, 1), 0.0, 0.5, transmission=1.0, thickness=1.0,
                              normal=(0, 0, -1), clearcoat_normal=(0, 0, -1)), False),
        ("clearcoat", ms((0.2, 0.3, 0.8), 0.0, 0.7, clearcoat=1.0, clearcoat_roughness=0.4), True),
        ("sheen", ms((0.3, 0.1, 0.1), 0.0, 0.8, sheen_color=(0.9, 0.7, 0.7), sheen_roughness=0.6), True),
    ]
    return cases


def _cases_ext():
    ms = A.material_sample
    brushed = dict(anisotropy=0.9, anisotropy_direction=(1, 0, 0))
    return [
        ("brushed metal (along v)", ms((0.95, 0.9, 0.8), 1.0, 0.5, **brushed), True),
        ("brushed metal (across v)", ms((0.95, 0.9, 0.8), 1.0, 0.5, anisotropy=0.9,
                                        anisotropy_direction=(0, 1, 0)), True),
        ("anisotropic dielectric", ms((0.7, 0.3, 0.2), 0.0, 0.45, anisotropy=0.7,
                                      anisotropy_direction=(1, 0, 0)), True),
        ("anisotropic thin transmission", ms((0.7, 0.3, 0.2), 0.0, 0.45, anisotropy=0.7,
                                             anisotropy_direction=(1, 0, 0), transmission=0.8), True),
        ("iridescent dielectric", ms((0.8, 0.8, 0.8), 0.0, 0.3, iridescence=1.0, iridescence_thickness=550.0), True),
        ("iridescent metal", ms((0.9, 0.6, 0.3), 1.0, 0.4, iridescence=0.8, iridescence_thickness=300.0), True),
        ("diffuse transmission", ms((0.3, 0.6, 0.2), 0.0, 1.0, diffuse_transmission=0.6,
                                    diffuse_transmission_color=(0.9, 0.5, 0.3)), True),
        ("diffuse + specular transmission", ms((0.3, 0.6, 0.2), 0.0, 0.5, diffuse_transmission=0.6,
                                               diffuse_transmission_color=(0.9, 0.5, 0.3), transmission=0.3), True),
    ]


@pytest.mark.parametrize("case", _cases_core() + _cases_ext(), ids=lambda c: c[0])
def test_bsdf_sampling_matches_quadrature_and_conserves_energy(oracle, case):
    # src/pathtrace.rs:2954-3065.
    name, ms, front = case
    v = _norm([0.6, 0.0, 0.8])
    ng = np.array([0, 0, 1.0 if front else -1.0], np.float32)
    eq, pdf_int = quadrature(oracle, ms, v, ng, front)
    em = monte_carlo(oracle, ms, v, ng, front, 1 << 16)
    for k in range(3):
        tol = 0.02 * max(eq[k], 0.05)
        assert abs(eq[k] - em[k]) < tol, (name, k, eq, em)
        assert eq[k] <= 1.01, (name, eq)
    assert 0.5 < pdf_int <= 1.01, (name, pdf_int)


def test_sampled_pdf_equals_evaluated_pdf(oracle):
    # src/pathtrace.rs:3111-3121: sample() reports eval()'s mixture pdf.
    ms = A.material_sample((0.9, 0.9, 0.9), 1.0, 0.4, anisotropy=0.8, anisotropy_direction=(1, 0, 0))
    v = _norm([0.6, 0.0, 0.8])
    u = oracle.sobol_owen_4d(np.arange(64, dtype=np.uint32), np.full(64, 7, np.uint32))[:, :3]
    l, f, pdf, ok = oracle.bsdf_sample(ms, v, [0, 0, 1], True, u)
    f2, pdf2 = oracle.bsdf_eval(ms, v, [0, 0, 1], True, l[ok])
    assert np.all(np.abs(pdf[ok] - pdf2) <= 1e-4 * np.maximum(pdf[ok], 1.0))
    assert np.array_equal(f[ok], f2)


def test_anisotropy_zero_is_isotropic_and_strength_stretches(oracle):
    # src/pathtrace.rs:3067-3110.
    v = _norm([0.6, 0.0, 0.8])
    ng = [0, 0, 1]
    base = A.material_sample((0.9, 0.9, 0.9), 1.0, 0.4)
    zero = A.material_sample((0.9, 0.9, 0.9), 1.0, 0.4, anisotropy_direction=(1, 0, 0))
    r = [-0.6, 0.0, 0.8]
    dirs = np.array([r, _norm([-0.3, 0.1, 0.9]), _norm([0.2, -0.5, 0.8]), _norm([-0.8, 0.3, 0.5])], np.float32)
    a = oracle.bsdf_eval(base, v, ng, True, dirs)
    b = oracle.bsdf_eval(zero, v, ng, True, dirs)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    lx = _norm(np.add(r, [0.35, 0, 0]))
    ly = _norm(np.add(r, [0, 0.35, 0]))

    def ratio(ms):
        f, _ = oracle.bsdf_eval(ms, v, ng, True, np.stack([lx, ly]))
        return f[0, 0] / f[1, 0]

    iso = ratio(base)
    ax = A.material_sample((0.9, 0.9, 0.9), 1.0, 0.4, anisotropy=0.8, anisotropy_direction=(1, 0, 0))
    ay = A.material_sample((0.9, 0.9, 0.9), 1.0, 0.4, anisotropy=0.8, anisotropy_direction=(0, 1, 0))
    assert ratio(ax) > 2.0 * iso
    assert ratio(ay) < 0.5 * iso


def test_iridescent_fresnel_bounds(oracle):
    # src/pathtrace.rs:3124-3165.
    f0 = [0.04] * 3
    f = oracle.iridescent_fresnel(1.0, 1.3, f0, 400.0, 1.0)
    assert ((0 <= f) & (f <= 1)).all()
    spread = lambda x: x.max() - x.min()
    assert spread(f) > 0.02
    g = oracle.iridescent_fresnel(1.0, 1.3, f0, 401.0, 1.0)
    assert np.all(np.abs(f - g) < 0.02)
    assert spread(oracle.iridescent_fresnel(1.0, 1.3, f0, 20000.0, 1.0)) < 2e-3
    assert (oracle.iridescent_fresnel(1.0, 1.3, f0, 400.0, 0.0) > 0.95).all()
    assert np.array_equal(oracle.iridescent_fresnel(1.5, 1.0, f0, 400.0, 0.1), [1.0, 1.0, 1.0])
    v = _norm([0.3, 0.0, 0.95])
    r = np.array([[-0.3, 0.0, 0.95]], np.float32)
    plain = A.material_sample((0.8, 0.8, 0.8), 0.0, 0.2)
    off = A.material_sample((0.8, 0.8, 0.8), 0.0, 0.2, iridescence_thickness=550.0)
    on = A.material_sample((0.8, 0.8, 0.8), 0.0, 0.2, iridescence_thickness=550.0, iridescence=1.0)
    fp, pp = oracle.bsdf_eval(plain, v, [0, 0, 1], True, r)
    fo, po = oracle.bsdf_eval(off, v, [0, 0, 1], True, r)
    assert np.array_equal(fp, fo) and np.array_equal(pp, po)
    fi, _ = oracle.bsdf_eval(on, v, [0, 0, 1], True, r)
    assert spread(fp[0]) / fp[0].max() < 1e-3
    assert spread(fi[0]) / fi[0].max() > 0.1


def test_diffuse_transmission_closed_forms(oracle):
    # src/pathtrace.rs:3167-3202.
    n = [0, 0, 1]
    ms = A.material_sample((1, 1, 1), 0.0, 1.0, specular=0.0, diffuse_transmission=1.0,
                           diffuse_transmission_color=(0.5, 1.0, 0.25))
    fa, pa = oracle.bsdf_eval(ms, n, n, True, _norm([0.3, 0.2, 0.9])[None])
    assert np.array_equal(fa[0], [0, 0, 0]) and pa[0] == 0
    l = _norm([0.3, 0.2, -0.9])
    fb, pb = oracle.bsdf_eval(ms, n, n, True, l[None])
    nl = -l[2]
    want = np.array([0.5, 1.0, 0.25]) * nl / math.pi
    assert np.all(np.abs(fb[0] - want) < 1e-5)
    assert abs(pb[0] - nl / math.pi) < 1e-5
    assert oracle.bsdf_probs(ms, n, n, True)[5] == 1.0
    e = monte_carlo(oracle, ms, n, n, True, 4096)
    assert np.all(np.abs(e - [0.5, 1.0, 0.25]) < 2e-3)
    ms2 = A.material_sample((1, 1, 1), 0.0, 1.0, specular=0.0, diffuse_transmission=0.5,
                            diffuse_transmission_color=(0.5, 1.0, 0.25))
    fa, _ = oracle.bsdf_eval(ms2, n, n, True, np.array([[0, 0, 1]], np.float32))
    assert abs(fa[0, 0] - 0.5 / math.pi) < 1e-5
    e = monte_carlo(oracle, ms2, n, n, True, 4096)
    assert abs(e[1] - 1.0) < 2e-3 and abs(e[0] - 0.75) < 2e-3


def test_white_lambert_albedo_is_exact(oracle):
    # src/pathtrace.rs:3213-3228.
    ms = A.material_sample((1, 1, 1), 0.0, 1.0, specular=0.0)
    e = monte_carlo(oracle, ms, [0, 0, 1], [0, 0, 1], True, 4096)
    assert np.all(np.abs(e - 1.0) < 1e-3)


def test_octant_triangle_subtends_half_pi(oracle):
    o = oracle.triangle_solid_angle([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1])
    assert abs(o - math.pi / 2) < 1e-5


def test_spherical_triangle_samples_are_uniform_in_solid_angle(oracle):
    # src/pathtrace.rs:3238-3296.
    p = np.array([0.1, -0.2, 0.0], np.float32)
    a, b, c = (np.array(x, np.float32) for x in ([-0.6, 0.3, 1.0], [1.2, -0.4, 0.8], [0.2, 1.1, 1.4]))
    omega = oracle.triangle_solid_angle(p, a, b, c)
    g = lambda d: (d[..., 0] + 0.3) ** 2 + d[..., 1]
    n = np.cross(b - a, c - a)
    area2 = np.linalg.norm(n)
    nn = n / area2
    steps = 600
    i, j = np.meshgrid(np.arange(steps), np.arange(steps), indexing="ij")
    mask = j < steps - i
    u = (i[mask] + 1 / 3) / steps
    v = (j[mask] + 1 / 3) / steps
    q = a + (b - a) * u[:, None] + (c - a) * v[:, None]
    to = q - p
    r2 = (to * to).sum(1)
    d = to / np.sqrt(r2)[:, None]
    dw = np.abs(d @ nn) / r2
    expect = (g(d) * dw).sum() / dw.sum()
    assert abs(dw.sum() * 0.5 * area2 / steps ** 2 * 2.0 / omega - 1.0) < 0.02
    m = 1 << 15
    uu = oracle.sobol_owen_4d(np.arange(m, dtype=np.uint32), np.full(m, 5, np.uint32))
    s = 0.0
    for k in range(m):
        dd = oracle.sample_spherical_triangle(p, a, b, c, float(uu[k, 0]), float(uu[k, 1]))
        assert dd is not None
        if k % 97 == 0:
            t = np.dot(a - p, n) / np.dot(dd, n)
            w = oracle.barycentric_of(p + dd * t, a, b, c)
            assert (w > -1e-3).all()
        s += float(g(dd))
    got = s / m
    assert abs(got - expect) < 0.01 * max(abs(expect), 0.1), (got, expect)


def test_sheen_lut_is_a_bounded_albedo(oracle):
    lut = oracle.sheen_lut().reshape(16, 32)
    assert ((lut >= 0) & (lut <= 1)).all()
    # Directional albedo of the white sheen lobe is below 1 and non-zero.
    assert lut[8].min() > 0.0
