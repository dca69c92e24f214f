"""The oracle's display resolve (PathTracer::image / hdr,
src/pathtrace.rs:2593-2734; ToneMap::apply, srgb_to_linear and
linear_to_srgb_byte, src/hdr.rs:44-129), pinned by the reference's own
resolve tests (src/hdr.rs:222-279) and by the IEC 61966-2-1 curve restated
here in float64 numpy, independently of the oracle's code."""
import numpy as np
import pytest

import oracle as O
from oxr import _abi as A


def iec_encode(c):
    """linear -> sRGB-encoded (IEC 61966-2-1), float64, clamped, NaN -> 0."""
    c = np.nan_to_num(np.asarray(c, np.float64), nan=0.0).clip(0.0, 1.0)
    return np.where(c <= 0.0031308, 12.92 * c, 1.055 * c ** (1.0 / 2.4) - 0.055)


def iec_decode(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def test_srgb_decode_is_the_iec_curve():
    c = np.linspace(0.0, 1.0, 4097, dtype=np.float32)
    np.testing.assert_allclose(O.srgb_to_linear(c), iec_decode(c), rtol=2e-6, atol=1e-9)


def test_srgb_round_trips_every_byte():
    """srgb_round_trips_every_byte (src/hdr.rs:226-232): the byte of the
    decoded LUT entry is the byte itself."""
    lut = O.srgb_to_linear(np.arange(256, dtype=np.float32) / np.float32(255.0))
    assert np.array_equal(O.linear_to_srgb_byte(lut), np.arange(256, dtype=np.uint8))


def test_threshold_encoder_matches_powf_curve():
    """threshold_encoder_matches_powf_curve (src/hdr.rs:234-244): within one
    byte of round(encode(c) * 255) on 20,001 points, -1 -> 0, 7 -> 255.
    The thresholds are the decoded byte midpoints, so the two agree except
    next to a midpoint: in practice on every point here."""
    c = (np.arange(20001) / 20000.0).astype(np.float32)
    want = np.floor(iec_encode(c) * 255.0 + 0.5).astype(np.int64)
    got = O.linear_to_srgb_byte(c).astype(np.int64)
    assert np.abs(want - got).max() <= 1
    assert (want == got).mean() > 0.999
    assert list(O.linear_to_srgb_byte([-1.0, 7.0, np.nan, np.inf, -np.inf])) == [0, 255, 0, 255, 0]


def test_clamp_is_identity_in_range():
    """clamp_is_identity_in_range (src/hdr.rs:246-250), NaN included."""
    out = O.tone_map([[0.25, 0.5, 1.0], [4.0, -1.0, np.nan]], A.OXR_TONEMAP_CLAMP)
    assert out[0].tolist() == [0.25, 0.5, 1.0]
    assert out[1].tolist() == [1.0, 0.0, 0.0]


def test_reinhard_compresses_and_is_monotone():
    """reinhard_compresses_and_is_monotone (src/hdr.rs:252-258): L = 1 maps
    to 0.5; on greys the operator is v / (1 + v) (Reinhard 2002, eq. 3) and
    a colour keeps its chromaticity."""
    a = O.tone_map([[1.0, 1.0, 1.0], [100.0, 100.0, 100.0]], A.OXR_TONEMAP_REINHARD)
    assert abs(a[0, 0] - 0.5) < 1e-5
    assert a[0, 0] < a[1, 0] < 1.0
    v = np.linspace(0.0, 20.0, 201, dtype=np.float32)
    g = O.tone_map(np.repeat(v[:, None], 3, 1), A.OXR_TONEMAP_REINHARD)[:, 0]
    np.testing.assert_allclose(g, v / (1.0 + v.astype(np.float64)), rtol=2e-6, atol=1e-7)
    assert np.all(np.diff(g) >= 0)
    c = O.tone_map([[0.2, 0.4, 0.1]], A.OXR_TONEMAP_REINHARD)[0]
    np.testing.assert_allclose(c / c[0], [1.0, 2.0, 0.5], rtol=1e-6)


def test_aces_fit_matches_published_points():
    """aces_fit_matches_published_points (src/hdr.rs:260-275): f(0) = 0,
    f(1000) within 1e-3 of 1, monotone on 0.1..9.9, mid-grey 0.18 in
    [0.1, 0.25); and the Narkowicz rational fit with the 0.6 pre-scale,
    evaluated here in float64."""
    assert O.tone_map([[0.0, 0.0, 0.0]], A.OXR_TONEMAP_ACES)[0, 0] == 0.0
    assert abs(O.tone_map([[1000.0] * 3], A.OXR_TONEMAP_ACES)[0, 0] - 1.0) < 1e-3
    x = (np.arange(1, 100) * 0.1).astype(np.float32)
    v = O.tone_map(np.repeat(x[:, None], 3, 1), A.OXR_TONEMAP_ACES)[:, 0]
    assert np.all(np.diff(v) >= 0)
    g = O.tone_map([[0.18] * 3], A.OXR_TONEMAP_ACES)[0, 0]
    assert 0.1 <= g < 0.25
    y = x.astype(np.float64) * 0.6
    fit = np.clip((y * (2.51 * y + 0.03)) / (y * (2.43 * y + 0.59) + 0.14), 0.0, 1.0)
    np.testing.assert_allclose(v, fit, rtol=5e-6, atol=1e-7)


def _accum(px, count, n=1):
    """An accumulator whose pixels hold mean `px` over `count` samples."""
    a = np.zeros((1, n, 4), np.float64)
    a[..., :3] = np.asarray(px, np.float64) * count
    a[..., 3] = count
    return a


def test_hdr_to_rgba8_and_back():
    """hdr_to_rgba8_and_back (src/hdr.rs:277-287) through the resolve: a
    fully covered pixel of mean (0.5, 0, 2) gives bytes (188, 0, 255, 255);
    exposure 0 gives (0, 0, 0, 255); the HDR resolve returns the mean."""
    acc = _accum([0.5, 0.0, 2.0], 4.0)
    assert O.resolve_rgba8(acc, 4)[0, 0].tolist() == [188, 0, 255, 255]
    assert O.resolve_rgba8(acc, 4, exposure=0.0)[0, 0].tolist() == [0, 0, 0, 255]
    np.testing.assert_array_equal(O.resolve_hdr(acc, 4)[0, 0], np.float32([0.5, 0.0, 2.0, 1.0]))
    back = iec_decode(188 / 255.0)
    assert abs(back - 0.5) < 0.01


def test_uncovered_pixels_keep_the_background_bytes():
    """PathTracer::image (src/pathtrace.rs:2694-2716): a pixel with no
    covered sample keeps the background bytes exactly; the HDR resolve
    returns the decoded background."""
    acc = np.zeros((1, 1, 4))
    bg = (10, 200, 33, 77)
    assert O.resolve_rgba8(acc, 5, background=bg)[0, 0].tolist() == list(bg)
    h = O.resolve_hdr(acc, 5, background=bg)[0, 0]
    np.testing.assert_allclose(h[:3], iec_decode(np.array(bg[:3]) / 255.0), rtol=2e-6)
    assert h[3] == pytest.approx(77 / 255.0)


def test_coverage_mix_and_alpha_rounds_half_away_from_zero():
    """PathTracer::mix (src/pathtrace.rs:2605-2622): with coverage f and a
    transparent background the colour is the covered mean and alpha is f.
    One covered sample of six gives alpha 255/6 = 42.5 in f32, which
    f32::round takes to 43 (half away from zero)."""
    acc = _accum([0.25, 0.5, 1.0], 1.0)
    out = O.resolve_rgba8(acc, 6, background=(0, 0, 0, 0))[0, 0]
    assert out[3] == 43
    assert out[:3].tolist() == O.linear_to_srgb_byte([0.25, 0.5, 1.0]).tolist()
    # An opaque background: colour = f * m + (1 - f) * bg, alpha 1.
    acc = _accum([1.0, 1.0, 1.0], 1.0)
    h = O.resolve_hdr(acc, 4, background=(0, 0, 0, 255))[0, 0]
    np.testing.assert_allclose(h, [0.25, 0.25, 0.25, 1.0], rtol=1e-6)
