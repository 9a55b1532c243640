"""Radio post step, tone map and HUD on the GPU (bh_post_radio, bh_tonemap, bh_tonemap_hud) against the oracle's
restatement (oracle/post.cpp): the same float operations in the same order, so the blurred frames
are bit-identical, the statistics equal and the 8-bit images byte-identical. Inputs: a traced
256^2 radio frame and seeded random frames of ragged sizes; every pyramid depth of the beam blur
(sigma at full resolution, 1-3 levels down) and every tone-map option."""
import numpy as np
import pytest

import bhrt_amd as bh
from helpers import setup, shot, tracer_for
from oracle_ctypes import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def radio():
    cfg = setup(shot("c2_radio_1024.ini"), W=256, H=256)
    cfg["S"].n0 = 291464.1043528089
    t = tracer_for(cfg)
    img, _ = t.render(cfg["F"], t.opts(precision=bh.PREC_FP64))
    yield t, img
    t.close()


@pytest.fixture(scope="module")
def orc():
    return Oracle("bhm")


def frames(img):
    rng = np.random.default_rng(5)
    out = [("traced 256^2", img / img[..., 0].max())]
    for H, W in [(97, 130), (31, 7), (2, 300), (1, 9), (513, 257)]:
        a = rng.random((H, W, 4)).astype(np.float32)
        a[..., 1:] *= 0.3
        out.append((f"random {H}x{W}", a))
    return out


@pytest.mark.parametrize("sigma", [0.4, 1.7, 3.0, 6.5, 13.0, 27.0])
def test_blur_and_stats_match_oracle(radio, orc, sigma):
    t, img = radio
    for name, a in frames(img):
        got, lp, lm = t.post_radio(a, blur_sigma_px=sigma, hide_direct=0.3, percentile=0.995)
        want = orc.post_blur(a, sigma)
        bad = np.argwhere((got.view(np.int32) != want.view(np.int32)).any(axis=-1))
        assert len(bad) == 0, f"{name} sigma {sigma}: {len(bad)} pixels differ, first {bad[:3].tolist()}"
        wlp, wlm = orc.frame_stats(want, 0.3, 0.995)
        assert (lp, lm) == (wlp, wlm), name


@pytest.mark.parametrize("opts", [
    dict(),
    dict(colormap=1, exposure=2.5),
    dict(gamma=0.6, exposure=3.0),
    dict(overlay=0.7, hide_direct=0.5, exposure=1.5),
    dict(vignette=0.15, grain=0.006, seed=0x51ED27 + 3 * 7919),
    dict(colormap=1, gamma=1.4, overlay=1.0, fade=0.4, vignette=0.3, grain=0.02, seed=99, exposure=4.0),
])
def test_tonemap_matches_oracle(radio, orc, opts):
    t, img = radio
    for name, a in frames(img):
        got = t.tonemap(a, **opts)
        want = orc.tonemap_radio(a, **opts)
        assert got.shape == want.shape
        nbad = int((got != want).any(axis=-1).sum())
        assert nbad == 0, f"{name} {opts}: {nbad} pixels differ"


def hud_case(rng, W, H, ng, nr, glyphs_in_font):
    glyphs = [(float(rng.uniform(-30, W)), float(rng.uniform(-30, H)), float(rng.uniform(0.2, 3.0)),
               int(rng.integers(0, glyphs_in_font)), int(rng.integers(0, 2 ** 32))) for _ in range(ng)]
    rects = [(int(rng.integers(-20, W)), int(rng.integers(-20, H)), int(rng.integers(-20, W + 20)),
              int(rng.integers(-20, H + 20)), int(rng.integers(0, 2 ** 32)), int(rng.integers(0, 3))) for _ in range(nr)]
    return glyphs, rects


def test_hud_matches_oracle(radio, orc):
    """bh_tonemap_hud against the oracle's tone map + HUD (render.cu:409-446, 811-846): random
    overlapping rectangles in three layers and glyphs of random scale and position (partly off the
    image) from a synthetic 99-glyph atlas, byte for byte; plus a title-like line of glyphs."""
    t, img = radio
    rng = np.random.default_rng(9)
    atlas = rng.integers(0, 256, (99, 44, 24), dtype=np.uint8)
    atlas[:, :4] = 0
    for name, a in frames(img):
        H, W = a.shape[:2]
        glyphs, rects = hud_case(rng, W, H, 300, 40, 99)
        glyphs += [(10.0 + 21.6 * 0.6 * i, 20.0, 0.6, (7 * i) % 99, 0xF0FFFFFF) for i in range(20)]
        opts = dict(exposure=2.0, vignette=0.1, grain=0.01, seed=17)
        got = t.tonemap(a, glyphs=glyphs, rects=rects, atlas=atlas, advance=21.6, baseline=34, **opts)
        want = orc.hud_draw(orc.tonemap_radio(a, **opts), glyphs, rects, atlas, max(1.5, H / 720.0))
        nbad = int((got != want).any(axis=-1).sum())
        assert nbad == 0, f"{name}: {nbad} pixels differ"
        plain = t.tonemap(a, **opts)
        assert (got != plain).any()
    # rectangles only (no font needed) and the reference's element limits (1024 rects, 4096 glyphs)
    a = frames(img)[0][1]
    _, rects = hud_case(rng, 256, 256, 0, 1100, 1)
    got = t.tonemap(a, rects=rects)
    want = orc.hud_draw(orc.tonemap_radio(a), [], rects[:1024], np.zeros((1, 1, 1), np.uint8), 1.5)
    assert np.array_equal(got, want)


def test_hud_errors(radio):
    t, img = radio
    atlas = np.zeros((10, 8, 6), np.uint8)
    with pytest.raises(bh.BhError, match="outside the font"):
        t.tonemap(img, glyphs=[(1.0, 1.0, 1.0, 10, 0xFFFFFFFF)], atlas=atlas)
    with pytest.raises(bh.BhError, match="scale must be"):
        t.tonemap(img, glyphs=[(1.0, 1.0, 0.0, 1, 0xFFFFFFFF)], atlas=atlas)
    with pytest.raises(bh.BhError, match="need a font"):
        t.tonemap(img, glyphs=[(1.0, 1.0, 1.0, 1, 0xFFFFFFFF)])


def test_post_errors(radio):
    t, img = radio
    with pytest.raises(bh.BhError, match="colormap"):
        t.tonemap(img, colormap=2)


def test_cli_writes_tone_mapped_frames(tmp_path):
    """The bhrt CLI renders a small radio shot with the beam blur and writes one PNG per frame."""
    import os
    import subprocess

    from helpers import ROOT, read_png
    exe = os.path.join(ROOT, "spacetime-optics_amd", "bhrt")
    sets = ["camera.width=64", "camera.height=64", "render.spp=1", "render.spp_max=1", "shot.frames=2",
            "post.blur=1", "flow.turbulence=0"]
    cmd = [exe, shot("c2_radio_1024.ini"), "--out", str(tmp_path)]
    for s_ in sets:
        cmd += ["--set", s_]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    for i in range(2):
        img = read_png(os.path.join(tmp_path, "frames", f"frame_{i:05d}.png"))
        assert img.shape == (64, 64, 3)
        assert img[..., 0].max() > 150   # exposure puts the 99.5th percentile near key 0.92
        assert img[0, 0].max() <= 4      # the corner stays dark (grain and dither only)
    assert os.path.exists(os.path.join(tmp_path, "meta.json"))


def test_cli_draws_hud_with_font_file(tmp_path):
    """bhrt --font FILE: the frames equal the ones without a font except inside the HUD's blocks
    (the layout of bhh_make_hud for the same frame), where text and bars are drawn."""
    import os
    import subprocess

    from helpers import ROOT, read_png
    rng = np.random.default_rng(12)
    atlas = np.zeros((99, 44, 24), dtype=np.uint8)
    # coverage only where a font's letters are (rows between cap height and baseline, inner columns),
    # as the layout's text boxes assume; a full random cell would spill past them
    atlas[:, 6:30, 4:18] = rng.integers(0, 256, (99, 24, 14), dtype=np.uint8)
    font = os.path.join(tmp_path, "font.bin")
    bh.save_font(font, atlas, advance=21.6, baseline=34)
    exe = os.path.join(ROOT, "spacetime-optics_amd", "bhrt")
    sets = ["camera.width=160", "camera.height=160", "render.spp=1", "render.spp_max=1", "shot.frames=1",
            "flow.turbulence=0", "hud.style=radio", "hud.title=M87*", "post.overlay=0.5"]
    imgs = []
    for extra in ([], ["--font", font]):
        out = os.path.join(tmp_path, "f" if extra else "n")
        cmd = [exe, shot("c2_radio_1024.ini"), "--out", out] + extra
        for s_ in sets:
            cmd += ["--set", s_]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        imgs.append(read_png(os.path.join(out, "frames", "frame_00000.png")))
    plain, hud = imgs
    changed = (plain != hud).any(axis=-1)
    assert changed.sum() > 50
    glyphs, rects = bh.make_hud(shot("c2_radio_1024.ini"), ";".join(sets), 0, 160, 160, 44, 21.6)
    inside = np.zeros_like(changed)
    for x0, y0, x1, y1, _, _ in rects:
        inside[max(y0, 0):max(y1, 0), max(x0, 0):max(x1, 0)] = True
    assert not (changed & ~inside).any()
