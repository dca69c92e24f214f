"""The converged reference of the glossy-caustic window that bench.py's
time-to-MSE leg measures against (tests/golden/caustic_window_reference.npz,
written by tests/golden/make_caustic_reference.py on the GPU: plain
sampling, f64 accumulation) checked against the oracle's own f64 render of
a crop of that window with another seed: per pixel and channel within
3 sigma, and the crop means within 3 sigma of each other. Both are the same
plain estimator, so sigma uses the per-sample variance pooled over both
runs: away from the caustic the floor is lit by rare paths through the
glass, whose large contributions a few thousand oracle samples mostly miss
(and then under-estimate their own variance too)."""
import os

import numpy as np
import pytest

from oxr import scenes

REF = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "caustic_window_reference.npz")


@pytest.mark.skipif(not os.path.exists(REF), reason="reference window not generated")
def test_caustic_window_reference_agrees_with_the_oracle(oracle):
    ref = np.load(REF)
    x0, y0, w, h = (int(v) for v in ref["window"])
    W = int(ref["frame"][0])
    s = scenes.glossy_caustic()
    os_ = oracle.OracleScene(s)
    # An 8 x 8 crop from the window's top rows, on the caustic's edge.
    cx, cy = x0 + w // 2 - 4, y0 + 4
    n = 4096
    N = int(ref["spp"])
    opts = oracle.options(W, W, spp=n, max_bounces=int(ref["max_bounces"]), seed=3, ambient=s.ambient)
    acc, sq = os_.render_rect(W, W, 0, n, cx, cx + 8, cy, cy + 8, opts=opts)
    assert (acc[..., 3] == n).all()
    mc = acc[..., :3] / n
    vc_s = np.maximum(sq / n - mc * mc, 0.0)  # per-sample variance, oracle
    mr = ref["mean"][cy - y0:cy - y0 + 8, cx - x0:cx - x0 + 8]
    vr_s = ref["var_mean"][cy - y0:cy - y0 + 8, cx - x0:cx - x0 + 8] * N  # per-sample variance, reference
    var = (N * vr_s + n * vc_s) / (N + n)
    z = np.abs(mr - mc) / np.sqrt(var / n + var / N + 1e-30)
    assert (z > 3).mean() < 0.02, ((z > 3).mean(), z.max())
    zm = abs(mr.mean() - mc.mean()) / np.sqrt((var / n + var / N).sum() / mr.size ** 2)
    assert zm < 3, zm
    assert mr.max() > 1.0  # the crop holds part of the caustic
