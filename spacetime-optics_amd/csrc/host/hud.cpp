/* hud.cpp - text overlay layout and the font file.

   The overlay's content and format are bhrt's frame format (blackholes/src/main.cpp:106-300): the
   same blocks (title card; title and description; live numbers; notes; photon-ring order legend;
   scale bar), strings, number formats, colours, scales and anchor positions. Its layout here is
   declarative: make_hud only describes blocks of lines and items as data, and one placement pass
   turns them into glyph runs, bars and translucent backings. */
#include "hud.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace bh {

namespace {

constexpr char FONT_MAGIC[8] = {'B', 'H', 'F', 'O', 'N', 'T', '1', '\n'};

// Colours (ABGR as the glyph rgba): white, dim and accent text, the photon-ring order tints
// (main.cpp:186-187).
constexpr uint32_t WHITE = 0xF0FFFFFFu, DIM = 0xC8D8DDE2u, ACCENT = 0xF070C8FFu;
constexpr uint32_t ORDER_COLOUR[5] = {0xF0FFFFFFu, 0xF0FFE640u, 0xF0F24DFFu, 0xF033D1FFu, 0xF04DFF8Cu};

// '|'-separated fields, trimmed, as std::getline splits them (a trailing separator adds no field)
std::vector<std::string> fields(const std::string& s) {
    std::vector<std::string> out;
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find('|', p);
        if (q == std::string::npos) q = s.size();
        out.push_back(trim(s.substr(p, q - p)));
        p = q + 1;
    }
    return out;
}

int text_length(const std::string& s) {
    int n = 0;
    for (size_t i = 0; i < s.size();) {
        glyph_of(s, &i);
        ++n;
    }
    return n;
}

/* The overlay is described as data, then placed. A Block is a stack of Lines from (x, y); each Line
   holds one or more Items (text at a horizontal offset from x, with its own alignment) and the
   advance to the next line in cell heights of the line's scale. A block may carry bars (pixel
   rectangles drawn over its backing) and its own opacity factor; every block except the title card
   gets a translucent backing around everything it drew. */
enum Align { LEFT = 0, RIGHT = 1, CENTRE = 2 };

struct Item {
    std::string text;
    float dx;
    Align align;
    uint32_t rgba;
};
struct Line {
    std::vector<Item> items;
    float scale;
    float lead;   // advance to the next line: cell height x scale x lead
};
struct Bar {
    int x0, y0, x1, y1;
};
struct Block {
    Block(float x0, float y0, std::vector<Line> l = {}) : x(x0), y(y0), lines(std::move(l)) {}
    float x, y;
    std::vector<Line> lines;
    std::vector<Bar> bars;
    float opacity = 1.f;
    bool backing = true;
};

Line line(std::string text, Align align, float scale, uint32_t rgba, float lead) {
    return Line{{Item{std::move(text), 0.f, align, rgba}}, scale, lead};
}

// Places the blocks: glyph runs (layer order: backings 0, bars 1) with colours scaled by the frame's
// HUD opacity and the block's own.
class Placer {
public:
    Placer(float alpha, int cell_h, float advance) : alpha_(alpha), cell_h_((float)cell_h), adv_(advance) {}

    void place(const Block& b, float pad) {
        a_ = alpha_ * b.opacity;
        box_ = {1e9f, 1e9f, -1e9f, -1e9f};
        const size_t bars_at = out.rects.size();
        for (const Bar& r : b.bars) {
            out.rects.push_back(bh_hud_rect{r.x0, r.y0, r.x1, r.y1, tint(0xE0FFFFFFu), 1});
            cover((float)r.x0, (float)r.y0, (float)r.x1, (float)r.y1);
        }
        float y = b.y;
        for (const Line& l : b.lines) {
            for (const Item& it : l.items) run(b.x + it.dx, y, l.scale, it.text, it.rgba, it.align);
            y += cell_h_ * l.scale * l.lead;
        }
        if (b.backing && box_[2] >= box_[0])
            out.rects.insert(out.rects.begin() + (long)bars_at,
                             bh_hud_rect{(int)(box_[0] - pad), (int)(box_[1] - pad), (int)(box_[2] + pad),
                                         (int)(box_[3] + pad), tint(0x55000000u), 0});
    }
    HudLayout out;

private:
    float alpha_, cell_h_, adv_, a_ = 1.f;
    std::array<float, 4> box_{};

    uint32_t tint(uint32_t rgba) const {
        const uint32_t a = (uint32_t)((float)((rgba >> 24) & 255u) * a_);
        return (rgba & 0x00ffffffu) | (a << 24);
    }
    void cover(float x0, float y0, float x1, float y1) {
        box_[0] = std::min(box_[0], x0);
        box_[1] = std::min(box_[1], y0);
        box_[2] = std::max(box_[2], x1);
        box_[3] = std::max(box_[3], y1);
    }
    // one run of text: its extent is the advance times the character count; the ink spans 0.12 to 0.95
    // of the cell height
    void run(float x, float y, float scale, const std::string& s, uint32_t rgba, Align align) {
        const float w = (float)text_length(s) * adv_ * scale;
        if (align == RIGHT) x -= w;
        if (align == CENTRE) x -= 0.5f * w;
        cover(x, y + 0.12f * cell_h_ * scale, x + w, y + 0.95f * cell_h_ * scale);
        for (size_t i = 0; i < s.size();) {
            const int g = glyph_of(s, &i);
            if (g) out.glyphs.push_back(bh_hud_glyph{x, y, scale, g, tint(rgba)});
            x += adv_ * scale;
        }
    }
};

}  // namespace

int glyph_of(const std::string& s, size_t* i) {
    const unsigned char c = (unsigned char)s[*i];
    if (c < 0x80) {
        ++*i;
        return c >= 32 && c < 127 ? c - 32 : 0;
    }
    const unsigned char d = *i + 1 < s.size() ? (unsigned char)s[*i + 1] : 0;
    *i += 2;
    if (c == 0xC2 && d == 0xB5) return 95;   // micro
    if (c == 0xC2 && d == 0xB0) return 96;   // degree
    if (c == 0xC3 && d == 0x97) return 97;   // times
    if (c == 0xC2 && d == 0xB1) return 98;   // plus-minus
    return 0;
}

std::string fmt_int(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.0f", std::fabs(v));
    const std::string digits = b;
    std::string o;
    const int n = (int)digits.size();
    for (int i = 0; i < n; ++i) {
        o += digits[i];
        if ((n - i - 1) % 3 == 0 && i != n - 1) o += ',';
    }
    return (v < 0 ? "-" : "") + o;
}

std::string fmt_g(double v, const char* f) {
    char b[64];
    std::snprintf(b, sizeof b, f, v);
    return b;
}

std::string fmt_uas(double uas) {
    if (uas >= 1000) return fmt_g(uas / 1000, "%.3g") + " mas";
    if (uas >= 0.9995) return fmt_g(uas, "%.3g") + " \xC2\xB5" "as";
    if (uas >= 1e-3) return fmt_g(uas * 1e3, "%.3g") + " nas";
    return fmt_g(uas * 1e6, "%.3g") + " pas";
}

bool load_font(const std::string& path, Font* font, std::string* err) {
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        *err = "cannot open font " + path;
        return false;
    }
    char magic[8];
    int32_t dims[4];
    float adv;
    const bool head = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, FONT_MAGIC, 8) == 0 &&
                      std::fread(dims, 4, 4, fp) == 4 && std::fread(&adv, 4, 1, fp) == 1;
    if (!head || dims[0] < 1 || dims[1] < 1 || dims[0] > 256 || dims[1] > 256 || dims[2] < 1 || dims[2] > 65536 ||
        !(adv > 0.f && adv < 1e4f)) {
        std::fclose(fp);
        *err = "not a font file (BHFONT1 header): " + path;
        return false;
    }
    font->cell_w = dims[0];
    font->cell_h = dims[1];
    font->glyphs = dims[2];
    font->baseline = dims[3];
    font->advance = adv;
    font->atlas.resize((size_t)dims[0] * dims[1] * dims[2]);
    const bool ok = std::fread(font->atlas.data(), 1, font->atlas.size(), fp) == font->atlas.size();
    std::fclose(fp);
    if (!ok) *err = "font file truncated: " + path;
    return ok;
}

bool save_font(const std::string& path, const Font& font, std::string* err) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) {
        *err = "cannot write " + path;
        return false;
    }
    const int32_t dims[4] = {font.cell_w, font.cell_h, font.glyphs, font.baseline};
    bool ok = std::fwrite(FONT_MAGIC, 1, 8, fp) == 8 && std::fwrite(dims, 4, 4, fp) == 4 &&
              std::fwrite(&font.advance, 4, 1, fp) == 1 &&
              std::fwrite(font.atlas.data(), 1, font.atlas.size(), fp) == font.atlas.size();
    ok = std::fclose(fp) == 0 && ok;
    if (!ok) *err = "cannot write " + path;
    return ok;
}

namespace {

// The numbers of the top-right block for the three HUD styles (main.cpp:231-271 decides the same
// content): radio - frequency, beam, elapsed time; zoom - magnification and field of view; visible
// - distance, orbital speed, inclination.
std::vector<Line> live_numbers(const Config& c, const std::string& style, const ShotFrame& f, const bh_frame& F,
                               const SceneInfo& si, const bh_scene& S, float big, float small, float lead_small) {
    std::vector<Line> L;
    const double width_m = f.fov_in_m ? f.fov / f.zoom : 0.0;
    if (style == "zoom") {
        L.push_back(line("zoom \xC3\x97" + fmt_int(f.zoom), RIGHT, big, WHITE, 1.05f));
        L.push_back(line("field of view " + fmt_uas(width_m * si.uas_per_m), RIGHT, small, ACCENT, lead_small));
        L.push_back(line("= " + fmt_g(width_m, "%.3g") + " M (gravitational radii)", RIGHT, small, DIM, lead_small));
    } else if (style == "radio") {
        L.push_back(line(fmt_g(S.nu_obs / 1e9, "%.0f") + " GHz", RIGHT, big, WHITE, 1.05f));
        const double beam = c.num("post.beam_uas", 20.0) * f.blur;
        L.push_back(line(beam > 0.5 ? "resolution " + fmt_g(beam, "%.0f") + " \xC2\xB5" "as (Event Horizon Telescope)"
                                    : std::string("resolution: unlimited"),
                         RIGHT, small, ACCENT, lead_small));
    } else {
        L.push_back(line("r = " + fmt_g(f.r, f.r < 100 ? "%.1f" : "%.0f") + " M", RIGHT, big, WHITE, 1.05f));
        const double km = f.r * si.rg_cm * 1e-5, au = km / 1.495978707e8;
        L.push_back(line((km < 1e8 ? fmt_int(km) + " km" : fmt_g(au, km < 1.5e10 ? "%.1f" : "%.0f") + " AU") +
                             " from the centre",
                         RIGHT, small, ACCENT, lead_small));
        const double v = std::sqrt(F.beta[0] * F.beta[0] + F.beta[1] * F.beta[1] + F.beta[2] * F.beta[2]);
        if (v > 1e-3) L.push_back(line("orbital speed " + fmt_g(v, "%.2f") + " c", RIGHT, small, DIM, lead_small));
        L.push_back(line("inclination " + fmt_g(f.incl * 180 / 3.14159265358979323846, "%.0f") + "\xC2\xB0", RIGHT,
                         small, DIM, lead_small));
        return L;
    }
    const double days = f.time_m * si.tg_s / 86400.0;
    L.push_back(line("elapsed " + fmt_g(days, days < 10 ? "%.1f" : "%.0f") + " days", RIGHT, small, DIM, lead_small));
    return L;
}

// Round angle of the scale bar: the largest 1, 2 or 5 times a power of ten not above a sixth of the
// image's angular width; its length in pixels
double round_angle(double pix_uas, int W, int* length_px) {
    const double target = pix_uas * W / 6.0, decade = std::pow(10.0, std::floor(std::log10(target)));
    double best = decade;
    for (double k : {2.0, 5.0, 10.0})
        if (k * decade <= target) best = k * decade;
    *length_px = (int)std::lround(best / pix_uas);
    return best;
}

}  // namespace

HudLayout make_hud(const Config& c, const Shot& shot, const ShotFrame& f, const bh_frame& F, const SceneInfo& si,
                   const bh_scene& S, int W, int H, int cell_h, float advance) {
    Placer P((float)(f.hud * f.fade), cell_h, advance);
    if ((float)(f.hud * f.fade) <= 0.01f) return P.out;
    // metrics in units of a 1080-line frame: margin, backing pad, three text scales, line lead
    const float u = (float)H / 1080.f, margin = 40 * u, pad = 10 * u;
    const float big = 0.85f * u, mid = 0.6f * u, small = 0.48f * u;
    const float lead = 1.12f, line_h = (float)cell_h * small * lead;
    const float fw = (float)W, fh = (float)H;

    std::vector<Block> blocks;
    if (shot.type == "title") {   // title card: up to six centred lines, no backing
        Block b{0.5f * fw, 0.f};
        b.backing = false;
        for (int i = 1; i <= 6; ++i) {
            const std::string key = "title.line" + std::to_string(i);
            if (!c.has(key)) continue;
            const bool first = b.lines.empty();
            b.lines.push_back(line(c.str(key), CENTRE, (first ? 1.6f : 0.62f) * u, first ? WHITE : DIM,
                                   first ? 1.3f : 1.25f));
        }
        b.y = 0.5f * fh - 0.5f * (float)b.lines.size() * (float)cell_h * 0.9f * u;
        blocks.push_back(std::move(b));
    } else {
        {   // top left: title and description
            Block b{margin, margin};
            if (c.has("hud.title")) b.lines.push_back(line(c.str("hud.title"), LEFT, mid, WHITE, 1.1f));
            for (const std::string& l : fields(c.str("hud.lines")))
                if (!l.empty()) b.lines.push_back(line(l, LEFT, small, DIM, lead));
            blocks.push_back(std::move(b));
        }
        // top right: live numbers
        blocks.push_back(Block{fw - margin, margin, live_numbers(c, c.str("hud.style", "visible"), f, F, si, S, big,
                                                                 small, lead)});
        {   // bottom left: notes, or the note that the direct image is hidden
            std::vector<std::string> notes = fields(c.str("hud.note"));
            if (f.hide_direct > 0.5) notes = {"direct image removed: only light that circled the black hole is shown"};
            if (!notes.empty() && !notes[0].empty()) {
                Block b{margin, fh - margin - line_h * (float)notes.size()};
                for (const std::string& l : notes) b.lines.push_back(line(l, LEFT, small, DIM, lead));
                blocks.push_back(std::move(b));
            }
        }
        if (f.overlay > 0.05) {   // bottom centre: photon-ring order legend while the order tint shows
            Block b{0.5f * fw, fh - margin - line_h * 3.2f};
            b.opacity = (float)std::min(1.0, f.overlay * 1.5);
            b.lines.push_back(line("photon-ring order n: half-orbits the light made around the hole", CENTRE, small, DIM,
                                   1.2f * lead * small / mid));   // the entries start 1.2 lines below
            Line entries{{}, mid, 1.f};   // one row, 150 px (at 1080 lines) apart, centred on the frame
            for (int n = 1; n <= 3; ++n)
                entries.items.push_back(Item{n < 3 ? "n = " + std::to_string(n) : "n = 3+", (float)(n - 2) * 150 * u - 75 * u,
                                             LEFT, ORDER_COLOUR[n]});
            b.lines.push_back(std::move(entries));
            blocks.push_back(std::move(b));
        }
        const double width_m = f.fov_in_m ? f.fov / f.zoom : 0.0;
        if (width_m > 0 && si.uas_per_m > 0) {   // bottom right: scale bar of a round angle
            int len = 0;
            const double ang = round_angle(width_m * si.uas_per_m / W, W, &len);
            const int x1 = (int)(fw - margin), x0 = x1 - len, yy = (int)(fh - margin - line_h * 0.4f);
            const int thick = std::max(2, (int)(3 * u)), tick = std::max(2, (int)(2 * u));
            Block b{0.5f * (float)(x0 + x1), (float)yy - line_h * 1.2f};
            b.bars = {{x0, yy, x1, yy + thick}, {x0, yy - (int)(8 * u), x0 + tick, yy + (int)(3 * u)},
                      {x1 - tick, yy - (int)(8 * u), x1, yy + (int)(3 * u)}};
            b.lines.push_back(line(fmt_uas(ang), CENTRE, small, WHITE, lead));
            blocks.push_back(std::move(b));
        }
    }
    for (const Block& b : blocks) P.place(b, pad);
    return P.out;
}

}  // namespace bh
