"""Pins of the guide oracle (oracle/guide_cpu.cpp) by mathematics, not by
restating it: finite differences of the loss, trilinear reproduction of
a linear field, the vMF mixture's normalisation, and the consistency of
the sampler with the pdf (E_q[g/q] = integral of g)."""
import math

import numpy as np
import pytest

import oracle as O
from oxr import _abi as A


def small_cfg(**kw):
    base = dict(log2_table=12, base_res=4, max_res=64, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0))
    base.update(kw)
    return A.guide_config(**base)


def init_params(cfg, seed=1, grid_scale=0.5):
    rng = np.random.default_rng(seed)
    grid = rng.uniform(-grid_scale, grid_scale, O.guide_num_grid(cfg)).astype(np.float32)
    mlp = np.zeros(A.GUIDE_MLP_PARAMS, np.float32)
    off = 0
    for nout, nin in [(64, 32), (64, 64), (64, 64), (48, 64)]:
        lim = math.sqrt(6.0 / (nin + nout))
        mlp[off:off + nout * nin] = rng.uniform(-lim, lim, nout * nin)
        off += nout * nin
        mlp[off:off + nout] = rng.uniform(-0.1, 0.1, nout)
        off += nout
    assert off == A.GUIDE_MLP_PARAMS
    return grid, mlp


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def batch(n, seed=2):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    dirs = unit(rng.normal(size=(n, 3)))
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    return pos, dirs, w


def test_finite_differences_pin_the_gradient():
    cfg = small_cfg()
    grid, mlp = init_params(cfg)
    pos, dirs, w = batch(6)
    loss, gg, gm = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=0)
    assert np.isfinite(loss) and loss > 0

    def fd(arr, i, h):
        a = arr.copy()
        a[i] = np.float32(arr[i] + h)
        hp = float(a[i]) - float(arr[i])
        lp = O.guide_loss_grad(cfg, a if arr is grid else grid, a if arr is mlp else mlp, pos, dirs, w, 0, False)[0]
        a[i] = np.float32(arr[i] - h)
        hm = float(arr[i]) - float(a[i])
        lm = O.guide_loss_grad(cfg, a if arr is grid else grid, a if arr is mlp else mlp, pos, dirs, w, 0, False)[0]
        return (lp - lm) / (hp + hm)

    rng = np.random.default_rng(5)
    # Every parameter block of the MLP, biases included.
    checked = 0
    for i in rng.choice(A.GUIDE_MLP_PARAMS, 60, replace=False):
        g_num = fd(mlp, i, 2e-3)
        assert abs(g_num - gm[i]) <= 2e-3 * max(abs(gm[i]), 1e-3), (i, g_num, gm[i])
        checked += abs(gm[i]) > 1e-6
    assert checked > 30
    touched = np.flatnonzero(gg)
    assert len(touched) > 100
    for i in rng.choice(touched, 40, replace=False):
        g_num = fd(grid, i, 2e-3)
        assert abs(g_num - gg[i]) <= 2e-3 * max(abs(gg[i]), 1e-3), (i, g_num, gg[i])


def test_device_mode_tracks_exact_mode():
    cfg = small_cfg()
    grid, mlp = init_params(cfg)
    pos, dirs, w = batch(64)
    f0, r0 = O.guide_forward(cfg, grid, mlp, pos, mode=0)
    f1, r1 = O.guide_forward(cfg, grid, mlp, pos, mode=1)
    assert np.abs(f1 - f0).max() <= 2 ** -8 * np.abs(f0).max()
    assert np.abs(r1 - r0).max() <= 0.03 * np.abs(r0).max()
    l0, g0, m0 = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=0)
    l1, g1, m1 = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=1)
    assert abs(l1 - l0) <= 0.01 * abs(l0)
    # The last layer's gradient carries one bf16 rounding (~0.4 %); deeper
    # ones inherit the cancellation of W^T dz (the logit gradient sums to
    # zero over lobes), which amplifies bf16 noise to several percent.
    nl = 64 * 32 + 64 + 2 * (64 * 64 + 64)
    assert np.linalg.norm(m1[nl:] - m0[nl:]) <= 0.02 * np.linalg.norm(m0[nl:])
    assert np.linalg.norm(m1 - m0) <= 0.15 * np.linalg.norm(m0)
    assert np.linalg.norm(g1 - g0) <= 0.15 * np.linalg.norm(g0)


def test_trilinear_reproduces_linear_field_on_dense_levels():
    cfg = small_cfg()
    res, dense = O.guide_resolutions(cfg)
    assert res[0] == 4 and res[-1] == 64 and np.all(np.diff(res.astype(int)) >= 0)
    assert dense[0] and not dense[-1]
    T = 1 << cfg.log2_table
    a, b, c, d = 0.3, -0.7, 0.2, 0.1
    grid = np.zeros((A.GUIDE_LEVELS, T, 2), np.float32)
    for lvl in np.flatnonzero(dense):
        r = int(res[lvl]) + 1
        i = np.arange(r ** 3)
        x, y, z = i % r, (i // r) % r, i // (r * r)
        grid[lvl, i, 0] = a * x / (r - 1) + b * y / (r - 1) + c * z / (r - 1) + d
        grid[lvl, i, 1] = 1.0
    pos, _, _ = batch(200)
    feat, _ = O.guide_forward(cfg, grid.ravel(), np.zeros(A.GUIDE_MLP_PARAMS, np.float32), pos, mode=0)
    xn = (pos + 1.0) / 2.0
    want = a * xn[:, 0] + b * xn[:, 1] + c * xn[:, 2] + d
    for lvl in np.flatnonzero(dense):
        np.testing.assert_allclose(feat[:, 2 * lvl], want, atol=2e-6)
        np.testing.assert_allclose(feat[:, 2 * lvl + 1], 1.0, atol=2e-6)
    idx, wts = O.guide_corners(cfg, pos)
    np.testing.assert_allclose(wts.sum(-1), 1.0, atol=1e-6)
    assert idx.max() < T
    # Positions outside the box clamp to its faces.
    feat_out, _ = O.guide_forward(cfg, grid.ravel(), np.zeros(A.GUIDE_MLP_PARAMS, np.float32),
                                  np.array([[5.0, -5.0, 0.0]], np.float32), mode=0)
    assert abs(feat_out[0, 0] - (a * 1.0 + b * 0.0 + c * 0.5 + d)) < 2e-6


def _one_lobe(mu, kappa):
    lob = np.zeros((1, A.GUIDE_LOBES, 5), np.float32)
    lob[0, :, 2] = 1.0
    lob[0, :, 3] = 1.0
    lob[0, 0, :3] = mu
    lob[0, 0, 3] = kappa
    lob[0, 0, 4] = 1.0
    return lob


@pytest.mark.parametrize("kappa", [0.05, 1.0, 10.0, 300.0, 2900.0])
def test_vmf_lobe_integrates_to_one(kappa):
    mu = unit([0.3, -0.5, 0.8])
    lob = _one_lobe(mu, kappa)
    # Integrate over the cosine to mu: 2 pi * int q(c) dc, with nodes dense
    # near c = 1 where a sharp lobe lives.
    t = np.linspace(0.0, 1.0, 20001)
    c = 1.0 - 2.0 * t ** 4
    u = unit(np.cross(mu, [1.0, 0.0, 0.0]))
    dirs = (np.outer(c, mu) + np.outer(np.sqrt(np.maximum(0.0, 1 - c * c)), u)).astype(np.float32)
    q = O.guide_pdf(np.repeat(lob, len(dirs), 0), dirs)
    integral = 2 * math.pi * np.trapezoid(q[::-1], c[::-1])
    assert abs(integral - 1.0) < 2e-3, integral


def test_sampler_matches_pdf():
    rng = np.random.default_rng(9)
    lob = np.zeros((1, A.GUIDE_LOBES, 5), np.float32)
    kap = [0.05, 2.0, 8.0, 30.0, 1.0, 5.0, 15.0, 0.5]
    pis = np.array([0.25, 0.1, 0.15, 0.1, 0.1, 0.1, 0.1, 0.1], np.float32)
    for k in range(A.GUIDE_LOBES):
        lob[0, k, :3] = unit(rng.normal(size=3))
        lob[0, k, 3] = kap[k]
        lob[0, k, 4] = pis[k]
    n = 200_000
    u = rng.random((n, 3)).astype(np.float32)
    lobes = np.repeat(lob, n, 0)
    d = O.guide_sample(lobes, u)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    q = O.guide_pdf(lobes, d)
    # E_q[g / q] = int g over the sphere: g = 1 -> 4 pi, g = x^2 -> 4 pi / 3.
    assert abs(np.mean(1.0 / q) / (4 * math.pi) - 1.0) < 0.02
    assert abs(np.mean(d[:, 0].astype(np.float64) ** 2 / q) / (4 * math.pi / 3) - 1.0) < 0.03


def test_adam_step_closed_form():
    # First step: m = (1 - b1) g, v = (1 - b2) g^2, bias-corrected ratio =
    # sign(g) (for eps -> 0); lazy skips exact zeros.
    p = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    g = np.array([0.5, -2.0, 0.0, 1e-3], np.float32)
    O.guide_adam(p, m, v, g, 1, 0.01, 0.9, 0.99, 1e-15, True)
    np.testing.assert_allclose(p, [0.99, 2.01, 3.0, 3.99], rtol=1e-6)
    assert m[2] == 0 and v[2] == 0
