"""Radio post step and display on the CPU: the oracle's restatement (oracle/post.cpp) pinned by what
the mathematics fixes, the PNG writer by the PNG format, the exposure smoothing by closed forms.

- Gaussian blur (render.cu:297-310): a constant image is unchanged; a point source at level 0 gives
  the separable normalised Gaussian kernel; the measured width of a blurred point follows the
  requested sigma through the pyramid levels (render.cu:744); a linear ramp is unchanged away from
  the edges (the box pyramid and the bilinear upsample keep linear functions, which pins their
  pixel-centre conventions).
- Tone map (render.cu:334-407): afmhot (2x, 2x - 1/2, 2x - 1) and the inferno nodes at known
  inputs, the dither is zero-mean with |d| < 1, fade 0 is black, a pixel lit only by n = 1 light
  takes the n = 1 tint.
- Statistics (render.cu:771-791): percentile and maximum of a known distribution, hide_direct.
"""
import os

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import read_png
from oracle_ctypes import Oracle


@pytest.fixture(scope="module")
def orc():
    return Oracle("bhm")


def test_blur_constant_image_unchanged(orc):
    # to rounding (the weighted sum over ~6 sigma taps and, through the pyramid, the bilinear weights);
    # exactly for 1.0 on the full-resolution level, where the tap sum and the weight sum coincide
    for H, W, s in [(64, 80, 2.0), (5, 300, 3.0), (33, 17, 9.0), (128, 128, 40.0), (5, 300, 7.0)]:
        img = np.full((H, W, 4), 0.37, dtype=np.float32)
        np.testing.assert_allclose(orc.post_blur(img, s), img, rtol=1e-6, atol=0)
        if s <= 6.0:
            one = np.ones((H, W, 4), dtype=np.float32)
            assert np.array_equal(orc.post_blur(one, s), one), (H, W, s)


def test_blur_point_source_level0_is_normalised_gaussian(orc):
    H = W = 61
    for s in (1.0, 2.5, 3.0):
        img = np.zeros((H, W, 4), dtype=np.float32)
        img[30, 30] = (1.0, 0.5, 0.25, 0.125)
        out = orc.post_blur(img, s)
        R = int(np.ceil(3.0 * np.float32(s)))
        i = np.arange(-R, R + 1)
        g = np.exp(-0.5 * i * i / (s * s))
        g /= g.sum()
        want = np.zeros((H, W))
        want[30 - R:30 + R + 1, 30 - R:30 + R + 1] = np.outer(g, g)
        for ch, amp in enumerate((1.0, 0.5, 0.25, 0.125)):
            np.testing.assert_allclose(out[..., ch], amp * want, rtol=2e-6, atol=1e-9)


@pytest.mark.parametrize("sigma", [2.0, 5.0, 7.0, 13.0, 25.0])
def test_blur_width_follows_sigma_through_pyramid(orc, sigma):
    H = W = 512
    img = np.zeros((H, W, 4), dtype=np.float32)
    img[255:257, 255:257, 0] = 1.0   # a 2 x 2 source centred on a pyramid cell corner
    out = orc.post_blur(img, sigma)[..., 0].astype(np.float64)
    assert abs(out.sum() - 4.0) < 1e-4 * 4.0
    y, x = np.mgrid[0:H, 0:W]
    cx = (out * x).sum() / out.sum()
    var = (out * (x - cx) ** 2).sum() / out.sum()
    # the Gaussian's variance (cut at 3 sigma: 0.9733 sigma^2) plus the 2 x 2 source (1/4) and the
    # box/bilinear smoothing of the pyramid level (at most a level-pixel^2); a sigma not divided by
    # the level's scale would come out 2^k times too wide
    k = 0
    while sigma / 2 ** (k + 1) > 3.0:
        k += 1
    lo, hi = 0.97 * sigma ** 2, sigma ** 2 + 0.25 + 1.0 * (2 ** k) ** 2
    assert lo <= var <= hi, (sigma, k, var, lo, hi)


def test_blur_keeps_linear_ramp_in_interior(orc):
    H, W = 256, 384
    y, x = np.mgrid[0:H, 0:W].astype(np.float32)
    ramp = 1.0 + 0.01 * x + 0.003 * y
    img = np.stack([ramp, 2 * ramp, 0 * ramp, ramp], axis=-1).astype(np.float32)
    for sigma in (2.0, 8.0, 20.0):
        out = orc.post_blur(img, sigma)
        m = int(3 * sigma) + 16
        np.testing.assert_allclose(out[m:-m, m:-m], img[m:-m, m:-m], rtol=2e-5)


def test_blur_below_half_pixel_is_identity(orc):
    img = np.random.default_rng(1).random((20, 30, 4)).astype(np.float32)
    assert np.array_equal(orc.post_blur(img, 0.5), img)


def test_frame_stats_closed_forms(orc):
    v = np.arange(1, 10001, dtype=np.float32)
    img = np.zeros((100, 100, 4), dtype=np.float32)
    img[..., 0] = v.reshape(100, 100)
    lp, lm = orc.frame_stats(img, 0.0, 0.995)
    assert lm == 10000.0
    assert lp == v[int(0.995 * 9999)]
    # hide_direct: I = 2 with 0.5 of it in n = 1 -> displayed 2 - h * 1.5
    one = np.array([[[2.0, 0.5, 0.0, 0.0]]], dtype=np.float32)
    for h in (0.0, 0.5, 1.0):
        lp, lm = orc.frame_stats(one, h, 0.5)
        assert abs(lm - (2.0 - h * 1.5)) < 1e-7


def tone(orc, x, **kw):
    img = np.zeros((64, 64, 4), dtype=np.float32)
    img[..., 0] = x
    return orc.tonemap_radio(img, **kw).astype(np.float64)


def test_tonemap_afmhot_and_inferno_known_points(orc):
    # dither d in (-1, 1): a byte is floor(255 c + 0.5 + d), within 1.5 of 255 c (clamped to [0, 255])
    for x in (0.0, 0.1, 0.3, 0.5, 0.75, 1.0):
        c = np.clip([2 * x, 2 * x - 0.5, 2 * x - 1.0], 0, 1)
        got = tone(orc, x)
        for k in range(3):
            assert np.all(np.abs(got[..., k] - 255 * c[k]) < 1.5), (x, k)
            assert abs(got[..., k].mean() - 255 * c[k]) < 0.08 or c[k] in (0.0, 1.0), (x, k)
    nodes = [(0.0015, 0.0005, 0.014), (0.087, 0.045, 0.224), (0.258, 0.039, 0.406), (0.416, 0.090, 0.433),
             (0.578, 0.148, 0.404), (0.735, 0.216, 0.330), (0.865, 0.317, 0.226), (0.955, 0.497, 0.086),
             (0.988, 0.998, 0.645)]
    for i, c in enumerate(nodes):
        got = tone(orc, i / 8.0, colormap=1)
        for k in range(3):
            assert np.all(np.abs(got[..., k] - 255 * c[k]) < 1.5), (i, k)


def test_tonemap_dither_zero_mean_and_fades(orc):
    got = tone(orc, 0.2)   # afmhot red = 0.4 -> 102
    assert abs(got[..., 0].mean() - 102.0) < 0.05
    assert got[..., 0].min() >= 101 and got[..., 0].max() <= 103
    black = tone(orc, 0.8, fade=0.0)
    assert black.max() <= 1.0
    burned = tone(orc, 3.0)   # x - 1 = 2 over: every channel burned to white (255.5 + d, clamped)
    assert burned.min() >= 254.0 and burned.mean() > 254.8


def test_tonemap_order_tint(orc):
    img = np.zeros((32, 32, 4), dtype=np.float32)
    img[..., 0] = 0.4
    img[..., 1] = 0.4   # all the light orbited once (n = 1)
    got = orc.tonemap_radio(img, overlay=1.0).astype(np.float64)
    c = np.array([0.8, 0.3, 0.0])   # afmhot(0.4)
    L = min(max((0.3 * c[0] + 0.59 * c[1] + 0.11 * c[2]) * 1.3 + 0.05, 0), 1)
    want = np.clip(L * np.array([0.25, 0.9, 1.0]), 0, 1)
    for k in range(3):
        assert np.all(np.abs(got[..., k] - 255 * want[k]) < 1.5), k


def test_png_writer_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    for shape in [(17, 23, 3), (4, 5, 4), (9, 1), (1, 40, 3)]:
        img = rng.integers(0, 256, size=shape, dtype=np.uint8)
        p = os.path.join(tmp_path, "x.png")
        bh.write_png(p, img)
        assert np.array_equal(read_png(p), img)
    with pytest.raises(bh.BhError):
        bh.write_png(os.path.join(tmp_path, "no", "such", "dir.png"), np.zeros((2, 2, 3), np.uint8))


def test_smooth_exposure_closed_forms():
    lp = np.full(40, 0.25)
    np.testing.assert_allclose(bh.smooth_exposure(lp, 0.92, 4.0), 0.92 / 0.25, rtol=1e-13)
    lp = np.array([1.0, 2.0, 4.0, 0.0, 8.0])
    np.testing.assert_allclose(bh.smooth_exposure(lp, 1.0, 0.0), [1.0, 0.5, 0.25, 1.0, 0.125], rtol=1e-13)
    # two frames one step apart, sigma 1: weights 1 and exp(-1/2) -> geometric mean with those weights
    lp = np.array([1.0, 4.0])
    w = np.exp(-0.5)
    want0 = 1.0 / np.exp((np.log(1.0) + w * np.log(4.0)) / (1 + w))
    assert abs(bh.smooth_exposure(lp, 1.0, 1.0)[0] - want0) < 1e-13
    assert np.all(bh.smooth_exposure(np.zeros(6), 0.9, 2.0) == 1.0)


def test_hud_rect_closed_forms(orc):
    """Rectangles (render.cu:434-446): alpha 255 paints the colour, alpha 0 leaves the image, partial
    alpha is d (1 - a) + s a rounded; clipped to the image; drawn by layer, list order within one."""
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (30, 40, 3), dtype=np.uint8)
    out = orc.hud_draw(img, [], [(5, 6, 15, 12, 0xFF204080, 0)], np.zeros((1, 4, 4), np.uint8), 1.5)
    assert np.all(out[6:12, 5:15] == [0x80, 0x40, 0x20]) and np.array_equal(out[:6], img[:6])
    out = orc.hud_draw(img, [], [(0, 0, 40, 30, 0x00FFFFFF, 0)], np.zeros((1, 4, 4), np.uint8), 1.5)
    assert np.array_equal(out, img)
    a = np.float32(0x55) / np.float32(255)
    out = orc.hud_draw(img, [], [(-10, -10, 100, 100, 0x55000000, 0)], np.zeros((1, 4, 4), np.uint8), 1.5)
    want = np.floor(img.astype(np.float32) * (np.float32(1) - a) + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(out, want)
    # layer 1 drawn after layer 0 whatever the list order
    two = [(0, 0, 10, 10, 0xFF0000FF, 1), (0, 0, 10, 10, 0xFF00FF00, 0)]
    out = orc.hud_draw(img, [], two, np.zeros((1, 4, 4), np.uint8), 1.5)
    assert np.all(out[:10, :10] == [255, 0, 0])


def test_hud_glyph_closed_forms(orc):
    """Glyphs (render.cu:409-432): at scale 1 and integer position the bilinear sample is the atlas
    coverage itself; white at full alpha gives d (1 - a) + 255 a; the shadow is the same coverage,
    black, 3/4 alpha, offset by `off`; a blank cell draws nothing."""
    cw, ch = 6, 9
    atlas = np.zeros((3, ch, cw), np.uint8)
    atlas[1] = np.random.default_rng(4).integers(0, 256, (ch, cw))
    img = np.full((40, 40, 3), 100, np.uint8)
    out = orc.hud_draw(img, [(10.0, 12.0, 1.0, 1, 0xFFFFFFFF)], [], atlas, 20.0)   # shadow far away
    a = atlas[1].astype(np.float32) / np.float32(255)
    want = np.floor(np.float32(100) * (1 - a) + np.float32(255) * a + np.float32(0.5)).astype(np.uint8)
    assert np.array_equal(out[12:12 + ch, 10:10 + cw, 0], want)
    am = np.float32(0.75)
    ws = np.floor(np.float32(100) * (1 - a * am) + np.float32(0.5)).astype(np.uint8)
    sh = out[32:32 + ch, 30:30 + cw, 0][:8, :10]   # the shadow's cell clipped by the image
    assert np.array_equal(sh, ws[:sh.shape[0], :sh.shape[1]])
    assert np.array_equal(orc.hud_draw(img, [(3.0, 4.0, 2.5, 0, 0xFFFFFFFF)], [], atlas, 1.5), img)
    # scale 2: the cell covers ceil(2 cw) + 1 pixels and samples between atlas texels
    big = orc.hud_draw(img, [(5.0, 5.0, 2.0, 1, 0xFFFFFFFF)], [], atlas, 40.0)
    changed = np.argwhere((big != img).any(axis=-1))
    assert changed[:, 0].max() < 5 + 2 * ch + 1 and changed[:, 1].max() < 5 + 2 * cw + 1
