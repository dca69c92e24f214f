"""The learned path guide on the GPU against its CPU oracle
(oracle/guide_cpu.cpp): MFMA inference, mixture pdf and sampling, the
loss gradient (hash grid + MLP), the Adam step; then, inside the path
tracer, that a guide with mix 0 changes nothing, that guided renders
converge to the plain estimator's image, and what training gains."""
import math

import numpy as np
import pytest

import oracle as O
from oxr import _abi as A

pytestmark = pytest.mark.gpu


def small_cfg(**kw):
    base = dict(log2_table=14, base_res=4, max_res=256, aabb_min=(-1.0, -1.0, -1.0), aabb_max=(1.0, 1.0, 1.0),
                seed=3)
    base.update(kw)
    return A.guide_config(**base)


def rand_params(cfg, seed=1, grid_scale=0.5):
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
    return grid, mlp


def unit(v):
    v = np.asarray(v, np.float64)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


@pytest.fixture(scope="module")
def guide_and_params():
    from oxr.guide import Guide
    cfg = small_cfg()
    g = Guide(0, config=cfg)
    grid, mlp = rand_params(cfg)
    g.set_params(grid, mlp)
    return cfg, g, grid, mlp


def test_params_roundtrip(guide_and_params):
    cfg, g, grid, mlp = guide_and_params
    gg, gm = g.params()
    assert np.array_equal(gg, grid) and np.array_equal(gm, mlp)


def test_inference_matches_oracle(guide_and_params):
    cfg, g, grid, mlp = guide_and_params
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1.1, 1.1, (4099, 3)).astype(np.float32)  # ragged: not a multiple of 16
    raw, lobes = g.infer(pos)
    _, want = O.guide_forward(cfg, grid, mlp, pos, mode=1)
    err = np.abs(raw - want)
    scale = np.abs(want).max()
    # Different f32 summation order inside the MFMA; rare bf16 rounding
    # flips of an activation move an output by ~2^-8 of its scale.
    assert np.median(err) <= 1e-5 * scale
    assert err.max() <= 2e-2 * scale
    want_lobes = O.guide_lobes(raw)
    np.testing.assert_allclose(lobes, want_lobes, rtol=1e-4, atol=1e-5)


def test_pdf_and_sampling_match_oracle(guide_and_params):
    cfg, g, grid, mlp = guide_and_params
    rng = np.random.default_rng(5)
    n = 2048
    pos = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    _, lobes = g.infer(pos)
    dirs = unit(rng.normal(size=(n, 3)))
    pdf = g.pdf(pos, dirs)
    np.testing.assert_allclose(pdf, O.guide_pdf(lobes, dirs), rtol=2e-4, atol=1e-7)
    u = rng.random((n, 3)).astype(np.float32)
    d, p = g.sample(pos, u)
    want = O.guide_sample(lobes, u)
    np.testing.assert_allclose(d, want, atol=2e-4)
    np.testing.assert_allclose(p, O.guide_pdf(lobes, d), rtol=2e-4, atol=1e-7)


def test_gradient_matches_oracle(guide_and_params):
    cfg, g, grid, mlp = guide_and_params
    rng = np.random.default_rng(6)
    n = 1000  # padded to 1024 with zero-weight samples on the device
    pos = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    dirs = unit(rng.normal(size=(n, 3)))
    w = rng.uniform(0.0, 2.0, n).astype(np.float32)
    loss, gg, gm = g.gradient(pos, dirs, w)
    l_o, gg_o, gm_o = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=1)
    assert abs(loss - l_o) <= 1e-4 * abs(l_o)
    assert np.linalg.norm(gm - gm_o) <= 2e-2 * np.linalg.norm(gm_o)
    assert np.linalg.norm(gg - gg_o) <= 2e-2 * np.linalg.norm(gg_o)
    # The last layer carries only one rounding: tight.
    nl = 64 * 32 + 64 + 2 * (64 * 64 + 64)
    assert np.linalg.norm(gm[nl:] - gm_o[nl:]) <= 2e-3 * np.linalg.norm(gm_o[nl:])
    # Entries the batch never touches stay exactly zero.
    assert np.array_equal(gg == 0, gg_o == 0)
    # A second call sees a clean gradient buffer.
    loss2, gg2, gm2 = g.gradient(pos, dirs, w)
    assert np.allclose(gg2, gg, rtol=1e-5, atol=1e-9) and np.allclose(gm2, gm, rtol=1e-5, atol=1e-9)


def test_gradient_of_a_concentrated_batch_matches_oracle(guide_and_params):
    """All samples inside one small region: the 16 samples of a lane group
    share their grid entries at every level, so nearly every contribution
    goes through the merge (the lowest lane of each set of equal entries
    adds the set's sum). Same bars as the spread-out batch."""
    cfg, g, grid, mlp = guide_and_params
    rng = np.random.default_rng(12)
    n = 1024
    pos = (np.array([0.31, -0.42, 0.17]) + rng.uniform(-2e-3, 2e-3, (n, 3))).astype(np.float32)
    dirs = unit(rng.normal(size=(n, 3)))
    w = rng.uniform(0.0, 2.0, n).astype(np.float32)
    loss, gg, gm = g.gradient(pos, dirs, w)
    l_o, gg_o, gm_o = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=1)
    assert abs(loss - l_o) <= 1e-4 * abs(l_o)
    assert np.linalg.norm(gm - gm_o) <= 2e-2 * np.linalg.norm(gm_o)
    assert np.linalg.norm(gg - gg_o) <= 2e-2 * np.linalg.norm(gg_o)
    assert np.array_equal(gg == 0, gg_o == 0)
    # Few entries carry the whole batch: their sums are large, and they match.
    top = np.argsort(-np.abs(gg_o))[:64]
    np.testing.assert_allclose(gg[top], gg_o[top], rtol=2e-2, atol=1e-3 * np.abs(gg_o).max())


def test_adam_step_matches_oracle():
    from oxr.guide import Guide
    cfg = small_cfg(lr=1e-3, eps=1e-8)
    g = Guide(0, config=cfg)
    grid, mlp = rand_params(cfg, seed=11)
    g.set_params(grid, mlp)
    rng = np.random.default_rng(7)
    n = 512
    pos = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    dirs = unit(rng.normal(size=(n, 3)))
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    _, gg, gm = O.guide_loss_grad(cfg, grid, mlp, pos, dirs, w, mode=1)
    g.train(pos, dirs, w, steps=1)
    ng, nm = g.params()
    # Oracle Adam on the oracle gradient; compare where the gradient is
    # well above the device/oracle difference (first step: +-lr).
    for p, gr, lazy, new in ((grid.copy(), gg, True, ng), (mlp.copy(), gm, False, nm)):
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        O.guide_adam(p, m, v, gr.astype(np.float32), 1, cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, lazy)
        sig = np.abs(gr) > 1e-3 * np.abs(gr).max()
        assert sig.sum() > 100
        np.testing.assert_allclose(new[sig], p[sig], atol=2e-6)
        if lazy:
            untouched = gr == 0
            assert np.array_equal(new[untouched], p[untouched])


def test_training_learns_a_direction_field():
    """Synthetic target: at x, radiance arrives from the direction of a
    fixed point, within a cone about 0.05 rad wide; the trained mixture
    concentrates on it. The bar: at the source direction the guide's
    density beats that of a uniform spread over a 30-degree cone,
    1 / (2 pi (1 - cos 30 deg)) = 1.19 sr^-1 (15x the uniform sphere's
    1 / (4 pi)): the target is six times narrower than that cone, so a
    mixture that has learned where the radiance comes from clears it,
    while one that spreads over a hemisphere (0.16 sr^-1) does not."""
    from oxr.guide import Guide
    cfg = small_cfg(lr=1e-2, seed=5)
    g = Guide(0, config=cfg)
    rng = np.random.default_rng(8)
    src = np.array([0.3, 0.9, -0.2], np.float32)
    n = 8192
    pos = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    to = unit(src - pos)
    # Directions: a tight cone around `to` (weight 1) plus uniform ones
    # (weight 0: no radiance).
    jitter = unit(to + 0.05 * rng.normal(size=(n, 3)))
    uni = unit(rng.normal(size=(n, 3)))
    P = np.concatenate([pos, pos])
    D = np.concatenate([jitter, uni])
    W = np.concatenate([np.ones(n), np.zeros(n)]).astype(np.float32)
    test = rng.uniform(-0.8, 0.8, (256, 3)).astype(np.float32)
    before = g.pdf(test, unit(src - test))
    l0 = g.train(P, D, W, steps=1)
    l1 = g.train(P, D, W, steps=300)
    after = g.pdf(test, unit(src - test))
    assert l1 < l0 - 2.0
    # The grid gradient is summed by float atomics, whose order differs
    # from run to run, and 300 Adam steps amplify that: over 13 runs the
    # ratio below spread over 16x .. 77x. 10x keeps it a coarse check;
    # the cone bar after it is the sharp one.
    assert np.median(after) > 10.0 * np.median(before)
    cone30 = 1.0 / (2.0 * math.pi * (1.0 - math.cos(math.radians(30.0))))
    assert np.median(after) > cone30, (np.median(after), cone30)


# ---------------------------------------------------------------------------
# Inside the tracer.
# ---------------------------------------------------------------------------
def _render(scene, W, H, spp, passes, guide=None, train=True, bounces=6, seed=0):
    import oxr
    pt = oxr.PathTracer(0)
    o = oxr.RenderOptions(width=W, height=H, ambient=scene.ambient)
    o.path_trace.samples_per_pixel = spp * passes
    o.path_trace.max_bounces = bounces
    o.path_trace.seed = seed
    pt.sync(scene, o)
    if guide is not None:
        pt.set_guide(guide, train=train)
    for _ in range(passes):
        pt.refine(spp)
    acc, sq = pt.accum(with_sum_sq=True)
    pt.close()
    return acc, sq


def test_mix_zero_guide_is_the_plain_estimator():
    from oxr import scenes as S
    from oxr.guide import Guide
    sc = S.glossy_caustic(segments=24, rings=16)
    plain, psq = _render(sc, 48, 40, 4, 2)
    g = Guide(0, mix=0.0, log2_table=16, steps=2, batch=4096)
    guided, gsq = _render(sc, 48, 40, 4, 2, guide=g)
    assert np.array_equal(plain, guided) and np.array_equal(psq, gsq)
    assert g.stats()["steps"] == 0  # mix 0: no eligible vertex, no records


def _mean_sigma(acc, sq):
    n = acc[..., 3:4]
    mean = acc[..., :3] / n
    var = np.maximum(sq / n - mean ** 2, 0.0) / np.maximum(n - 1, 1)
    return mean, var


def test_guided_render_converges_to_plain():
    from oxr import scenes as S
    from oxr.guide import Guide
    sc = S.glossy_caustic(segments=32, rings=24)
    W = H = 64
    plain_acc, plain_sq = _render(sc, W, H, 64, 8, seed=1)
    g = Guide(0, log2_table=18, batch=16384, steps=8, record_shift=0)
    guided_acc, guided_sq = _render(sc, W, H, 64, 8, guide=g, seed=2)
    assert g.stats()["steps"] == 64 and g.stats()["collected"] > 0
    mp, vp = _mean_sigma(plain_acc, plain_sq)
    mg, vg = _mean_sigma(guided_acc, guided_sq)
    # Unbiased: image means agree, and so do the means of 8x8-pixel tiles
    # (z-scores of tile means: 64 pixels each, closer to normal than single
    # pixels of a caustic, whose samples are Bernoulli-like).
    tot = (mg.mean() - mp.mean()) / math.sqrt((vp.sum() + vg.sum()) / mp.size ** 2)
    assert abs(tot) < 4.0, tot
    tile = lambda a: a.reshape(H // 8, 8, W // 8, 8, 3).sum((1, 3))  # noqa: E731
    zt = (tile(mg) - tile(mp)) / np.sqrt(tile(vp) + tile(vg) + 1e-30)
    assert np.abs(zt).max() < 5.0, np.abs(zt).max()
    assert np.mean(np.abs(zt) > 3.0) < 0.05


def _pass_variance(pt, passes, spp):
    """Mean per-pixel single-sample variance over `passes` further passes
    (from the f64 sums and sums of squares of each pass)."""
    prev_a, prev_s = pt.accum(with_sum_sq=True)
    out = []
    for _ in range(passes):
        pt.refine(spp)
        a, s = pt.accum(with_sum_sq=True)
        da, ds = a - prev_a, s - prev_s
        prev_a, prev_s = a, s
        n = np.maximum(da[..., 3:4], 1.0)
        m = da[..., :3] / n
        out.append(float((np.maximum(ds / n - m * m, 0.0) * n / np.maximum(n - 1.0, 1.0)).mean()))
    return float(np.mean(out)), prev_a, prev_s


def test_online_guide_reduces_caustic_variance():
    """The gain of the trained guide on the glossy-caustic scene: after
    online training (stopped at its step budget), the guided estimator's
    per-sample variance is well below the plain estimator's, and the two
    images agree. Most of the variance is the caustic under the sphere,
    reached through a small set of directions that the guide learns; the
    guided paths there carry a small throughput by design, so this also
    checks that roulette does not end them."""
    import oxr
    from oxr import scenes as S
    from oxr.guide import Guide
    sc = S.glossy_caustic()
    W = 256
    res = {}
    for guided in (False, True):
        pt = oxr.PathTracer(0)
        o = oxr.RenderOptions(width=W, height=W, ambient=sc.ambient)
        o.path_trace.samples_per_pixel = 1 << 20
        o.path_trace.max_bounces = 8
        o.path_trace.seed = 11 if guided else 12
        pt.sync(sc, o)
        g = None
        if guided:
            g = Guide(0, max_steps=96, record_shift=1)
            pt.set_guide(g, train=True)
            for _ in range(24):
                pt.refine(4)
            assert g.stats()["steps"] == 96
            pt.reset()
        var, acc, sq = _pass_variance(pt, 48, 4)
        if g is not None:
            assert g.stats()["steps"] == 96  # training stopped at its budget
        res[guided] = (var, acc, sq)
        pt.close()
    ratio = res[True][0] / res[False][0]
    assert ratio < 0.65, ratio
    mp, vp = _mean_sigma(res[False][1], res[False][2])
    mg, vg = _mean_sigma(res[True][1], res[True][2])
    tot = (mg.mean() - mp.mean()) / math.sqrt((vp.sum() + vg.sum()) / mp.size ** 2)
    assert abs(tot) < 4.0, tot


def test_collected_samples_carry_the_mixture_pdf():
    """The online samples of a pass: finite positions on the scene, unit
    directions, non-negative weights, and a proposal pdf of at least
    mix * q(direction) under the guide that drew them."""
    import oxr
    from oxr import scenes as S
    from oxr.guide import Guide
    sc = S.glossy_caustic(segments=32, rings=24)
    W = 96
    pt = oxr.PathTracer(0)
    o = oxr.RenderOptions(width=W, height=W, ambient=sc.ambient)
    o.path_trace.samples_per_pixel = 1 << 20
    o.path_trace.max_bounces = 6
    pt.sync(sc, o)
    box = dict(aabb_min=(-5.0, 0.0, -5.0), aabb_max=(5.0, 5.0, 5.0))
    g = Guide(0, record_shift=0, batch=8192, **box)
    pt.set_guide(g, train=True)
    for _ in range(6):
        pt.refine(4)
    grid, mlp = g.params()
    frozen = Guide(0, **box)
    frozen.set_params(grid, mlp)
    h = Guide(0, record_shift=0, batch=8192, **box)
    h.set_params(grid, mlp)
    pt.set_guide(h, train=True)
    pt.refine(4)
    pos, w, d, pdf = h.collected()
    assert len(pos) > 1000
    assert np.isfinite(pos).all() and (pos[:, 1] > -1e-3).all() and (np.abs(pos) < 5.01).all()
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-4)
    assert (w >= 0).all() and np.isfinite(w).all() and (pdf > 0).all()
    q = frozen.pdf(pos, d)
    assert (pdf >= 0.5 * q * (1.0 - 1e-3) - 1e-6).mean() > 0.999
    pt.close()
