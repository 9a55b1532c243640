/* hud.h - the text overlay of bhrt's radio and visible frames (make_hud and struct Hud,
   blackholes/src/main.cpp:106-300): layout on the host, drawing by bh_tonemap_hud.

   The font is a file (load_font): "BHFONT1\n", then int32 cell_w, cell_h, glyphs, baseline, float32
   advance, then glyphs x cell_h x cell_w coverage bytes (little endian). Glyph g is ASCII 32 + g for
   g < 95; 95..98 are U+00B5, U+00B0, U+00D7, U+00B1 (the reference's FONT_ATLAS order,
   common/font_atlas.h). */
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/bhrt.h"
#include "config.h"
#include "scene.h"
#include "shot.h"

namespace bh {

struct Font {
    std::vector<uint8_t> atlas;
    int cell_w = 0, cell_h = 0, glyphs = 0, baseline = 0;
    float advance = 0.f;
    bh_font view() const { return bh_font{atlas.data(), cell_w, cell_h, glyphs, advance, baseline}; }
};

bool load_font(const std::string& path, Font* font, std::string* err);
bool save_font(const std::string& path, const Font& font, std::string* err);

// Glyph index of the UTF-8 character at s[*i] (advances *i); 0 (space) for characters outside the font.
int glyph_of(const std::string& s, size_t* i);

struct HudLayout {
    std::vector<bh_hud_glyph> glyphs;
    std::vector<bh_hud_rect> rects;
};

// make_hud (main.cpp:195-300): title block, live numbers (frequency, beam, elapsed time; zoom and
// field of view; or distance, speed and inclination), notes, photon-ring order legend and scale
// bar for one frame of a shot, in a W x H image, laid out with the font's cell height and advance.
HudLayout make_hud(const Config& c, const Shot& shot, const ShotFrame& f, const bh_frame& F, const SceneInfo& si,
                   const bh_scene& S, int W, int H, int cell_h, float advance);

// Number formats of the overlay (main.cpp:106-128)
std::string fmt_int(double v);
std::string fmt_g(double v, const char* f = "%.3g");
std::string fmt_uas(double uas);

}  // namespace bh
