"""HUD layout on the host (make_hud, blackholes/src/main.cpp:106-300, restated in
spacetime-optics_amd/csrc/host/hud.cpp) and the font file. The reference's layout runs only inside
its CLI (main.cpp, which needs the CUDA renderer), so its own glyph lists cannot be produced here:
these checks pin what the layout rules fix - the strings of each block, right alignment at W - m,
the line spacing, the scale bar's round length and pixel size, the legend and note switches, the
centred title card, the alpha of fading frames - on the radio shot at 1920 x 1080 with the
reference's font metrics (cell height 44 px, advance 21.6 px)."""
import math
import os
import struct

import numpy as np
import pytest

import bhrt_amd as bh
from helpers import ROOT

INI = os.path.join(ROOT, "spacetime-optics_amd", "shots", "c2_radio_1024.ini")
W, H, CH, ADV = 1920, 1080, 44, 21.6
S = H / 1080.0
M = 40 * S


def text(glyphs):
    """Characters of a run of glyphs (95..98: micro, degree, times, plus-minus)."""
    extra = {95: "µ", 96: "°", 97: "×", 98: "±"}
    return "".join(extra.get(g[3], chr(32 + g[3])) for g in glyphs)


def lines(glyphs):
    """Glyph runs grouped by (y, scale): {(y, scale): (text with spaces restored, x0, x_end)}."""
    out = {}
    for g in glyphs:
        out.setdefault((round(g[1], 3), round(g[2], 4)), []).append(g)
    res = {}
    for k, gs in out.items():
        gs.sort(key=lambda g: g[0])
        step = ADV * gs[0][2]
        s, x = "", gs[0][0]
        for g in gs:
            n = int(round((g[0] - x) / step))
            s += " " * max(n - 1, 0) if s else ""
            s += text([g])
            x = g[0]
        res[k] = (s, gs[0][0], gs[-1][0] + step)
    return res


def test_radio_block_strings_and_alignment():
    g, r = bh.make_hud(INI, "hud.style=radio;hud.title=M87*;hud.lines=one line|second line;post.blur=1",
                       0, W, H, CH, ADV)
    L = lines(g)
    strings = {v[0]: v for v in L.values()}
    assert "M87*" in strings and strings["M87*"][1] == pytest.approx(M)
    assert "one line" in strings and "second line" in strings
    assert "230 GHz" in strings and "resolution 20 µas (Event Horizon Telescope)" in strings
    for k in ("230 GHz", "resolution 20 µas (Event Horizon Telescope)"):
        assert strings[k][2] == pytest.approx(W - M, abs=1e-2)   # right-aligned at W - m (float sums)
    assert any(s.startswith("elapsed ") and s.endswith(" days") for s in strings)
    # the title line (scale mid) then the description lines spaced by 44 * small * 1.12
    ys = sorted(y for (y, sc) in L if abs(sc - 0.48 * S) < 1e-4 and L[(y, sc)][1] == pytest.approx(M))
    assert ys[0] == pytest.approx(M + CH * 0.6 * S * 1.1, abs=1e-3)
    assert ys[1] - ys[0] == pytest.approx(CH * 0.48 * S * 1.12, abs=1e-3)
    # one translucent backing per block (layer 0) and the scale bar's three bars (layer 1)
    assert sum(1 for x in r if x[5] == 0) == 3 and sum(1 for x in r if x[5] == 1) == 3


def test_scale_bar_is_a_round_angle():
    g, r = bh.make_hud(INI, "hud.style=radio", 0, W, H, CH, ADV)
    bars = [x for x in r if x[5] == 1]
    main = max(bars, key=lambda x: x[2] - x[0])
    assert main[2] == int(W - M)
    label = [s for s in lines(g) if lines(g)[s][0].endswith("as")]
    assert label, "scale-bar label"
    s = lines(g)[label[0]][0]
    value = float(s.split()[0])
    lead = value / 10 ** math.floor(math.log10(value))
    assert round(lead, 6) in (1.0, 2.0, 5.0)
    # bar length = the round angle over the pixel scale; the angle is at most a sixth of the width
    bar = main[2] - main[0]
    assert W / 6 / 5 <= bar <= W / 6 + 1


def test_legend_notes_title_and_fade():
    g, _ = bh.make_hud(INI, "hud.style=radio;post.overlay=1", 0, W, H, CH, ADV)
    s = {v[0] for v in lines(g).values()}
    legend = [x for x in s if x.startswith("n = 1")]
    assert legend and "n = 2" in legend[0] and legend[0].endswith("n = 3+")   # one row, 150 px apart
    assert "photon-ring order n: half-orbits the light made around the hole" in s
    g, _ = bh.make_hud(INI, "hud.style=radio;post.hide_direct=1", 0, W, H, CH, ADV)
    assert "direct image removed: only light that circled the black hole is shown" in {v[0] for v in lines(g).values()}
    g, _ = bh.make_hud(INI, "shot.type=title;title.line1=Kerr;title.line2=230 GHz", 0, W, H, CH, ADV)
    L = lines(g)
    for s_, x0, x1 in L.values():
        assert (x0 + x1) / 2 == pytest.approx(W / 2, abs=1e-2)   # centred
    assert {v[0] for v in L.values()} == {"Kerr", "230 GHz"}
    g, r = bh.make_hud(INI, "hud.style=radio;hud.alpha=0", 0, W, H, CH, ADV)
    assert g == [] and r == []
    g, _ = bh.make_hud(INI, "hud.style=radio;hud.alpha=0.5", 0, W, H, CH, ADV)
    alphas = {x[4] >> 24 for x in g}
    assert alphas <= {int(0xF0 * 0.5), int(0xC8 * 0.5)}


def test_zoom_and_visible_styles():
    g, _ = bh.make_hud(INI, "hud.style=zoom;camera.zoom=1000", 0, W, H, CH, ADV)
    s = {v[0] for v in lines(g).values()}
    assert "zoom ×1,000" in s
    assert any(x.startswith("field of view ") for x in s)
    g, _ = bh.make_hud(INI, "hud.style=visible", 0, W, H, CH, ADV)
    s = {v[0] for v in lines(g).values()}
    assert "r = 1000 M" in s and "inclination 17°" in s
    assert any(x.endswith(" AU from the centre") or x.endswith(" km from the centre") for x in s)


def test_font_file_and_header_converter(tmp_path):
    rng = np.random.default_rng(2)
    atlas = rng.integers(0, 256, (99, 6, 5), dtype=np.uint8)
    body = ",".join(str(v) for v in atlas.ravel())
    hdr = tmp_path / "font.h"
    hdr.write_text("#pragma once\nconstexpr int FONT_CELL_W = 5;\nconstexpr int FONT_CELL_H = 6;\n"
                   "constexpr int FONT_GLYPHS = 99;\nconstexpr int FONT_BASELINE = 4;\n"
                   f"static const uint8_t FONT_ATLAS[FONT_GLYPHS * FONT_CELL_H * FONT_CELL_W] = {{\n{body}\n}};\n")
    got, base = bh.font_from_header(str(hdr))
    assert np.array_equal(got, atlas) and base == 4
    path = tmp_path / "font.bin"
    bh.save_font(str(path), got, baseline=base)
    raw = path.read_bytes()
    assert raw[:8] == b"BHFONT1\n"
    cw, ch, n, bl = struct.unpack("<4i", raw[8:24])
    adv, = struct.unpack("<f", raw[24:28])
    assert (cw, ch, n, bl) == (5, 6, 99, 4) and adv == pytest.approx(4.5)
    assert np.array_equal(np.frombuffer(raw[28:], np.uint8).reshape(99, 6, 5), atlas)
    with pytest.raises(bh.BhError):
        bh.make_hud(INI, "", 10 ** 6, W, H, CH, ADV)
