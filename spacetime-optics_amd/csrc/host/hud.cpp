/* hud.cpp - text overlay layout (blackholes/src/main.cpp:106-300) and the font file. */
#include "hud.h"

#include <algorithm>
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

// Glyph runs and rectangles with the bounding box of the current block (struct Hud, main.cpp:130-184)
class Layout {
public:
    Layout(float alpha, int cell_h, float advance) : alpha_(alpha), cell_h_(cell_h), adv_(advance) {}
    HudLayout out;
    float alpha_;

    uint32_t colour(uint32_t rgba) const {
        const uint32_t a = (uint32_t)((float)((rgba >> 24) & 255u) * alpha_);
        return (rgba & 0x00ffffffu) | (a << 24);
    }
    // align: 0 left, 1 right, 2 centre
    void text(float x, float y, float scale, const std::string& s, uint32_t rgba, int align = 0) {
        const float w = (float)text_length(s) * adv_ * scale;
        if (align == 1) x -= w;
        if (align == 2) x -= 0.5f * w;
        grow(x, y + 0.12f * (float)cell_h_ * scale, x + w, y + 0.95f * (float)cell_h_ * scale);
        for (size_t i = 0; i < s.size();) {
            const int g = glyph_of(s, &i);
            if (g) out.glyphs.push_back(bh_hud_glyph{x, y, scale, g, colour(rgba)});
            x += adv_ * scale;
        }
    }
    void rect(int x0, int y0, int x1, int y1, uint32_t rgba) {
        out.rects.push_back(bh_hud_rect{x0, y0, x1, y1, colour(rgba), 1});
        grow((float)x0, (float)y0, (float)x1, (float)y1);
    }
    void begin_block() { bx0_ = by0_ = 1e9f, bx1_ = by1_ = -1e9f; }
    // the block's translucent backing (layer 0, under its text and bars)
    void end_block(float pad) {
        if (bx1_ < bx0_) return;
        out.rects.push_back(bh_hud_rect{(int)(bx0_ - pad), (int)(by0_ - pad), (int)(bx1_ + pad), (int)(by1_ + pad),
                                        colour(0x55000000u), 0});
    }

private:
    int cell_h_;
    float adv_;
    float bx0_ = 1e9f, by0_ = 1e9f, bx1_ = -1e9f, by1_ = -1e9f;
    void grow(float x0, float y0, float x1, float y1) {
        bx0_ = std::min(bx0_, x0);
        by0_ = std::min(by0_, y0);
        bx1_ = std::max(bx1_, x1);
        by1_ = std::max(by1_, y1);
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

HudLayout make_hud(const Config& c, const Shot& shot, const ShotFrame& f, const bh_frame& F, const SceneInfo& si,
                   const bh_scene& S, int W, int H, int cell_h, float advance) {
    Layout h((float)(f.hud * f.fade), cell_h, advance);
    if (h.alpha_ <= 0.01f) return h.out;
    const float ch = (float)cell_h;
    const float s = (float)H / 1080.f, m = 40 * s, pad = 10 * s;
    const float big = 0.85f * s, mid = 0.6f * s, small = 0.48f * s;
    const float lh = ch * small * 1.12f;
    const std::string style = c.str("hud.style", "visible");

    if (shot.type == "title") {   // centred title card: up to six lines
        std::vector<std::string> lines;
        for (int i = 1; i <= 6; ++i) {
            const std::string k = "title.line" + std::to_string(i);
            if (c.has(k)) lines.push_back(c.str(k));
        }
        float y = 0.5f * (float)H - 0.5f * (float)lines.size() * ch * 0.9f * s;
        for (size_t i = 0; i < lines.size(); ++i) {
            const float sc = (i == 0 ? 1.6f : 0.62f) * s;
            h.text(0.5f * (float)W, y, sc, lines[i], i == 0 ? WHITE : DIM, 2);
            y += ch * sc * (i == 0 ? 1.3f : 1.25f);
        }
        return h.out;
    }
    // top left: title and description lines
    h.begin_block();
    float y = m;
    if (c.has("hud.title")) {
        h.text(m, y, mid, c.str("hud.title"), WHITE);
        y += ch * mid * 1.1f;
    }
    for (const std::string& l : fields(c.str("hud.lines")))
        if (!l.empty()) {
            h.text(m, y, small, l, DIM);
            y += lh;
        }
    h.end_block(pad);

    // top right: live numbers
    h.begin_block();
    y = m;
    const double uas_m = si.uas_per_m;
    const double width_m = f.fov_in_m ? f.fov / f.zoom : 0.0;
    const float xr = (float)W - m;
    if (style == "radio" || style == "zoom") {
        if (style == "zoom") {
            h.text(xr, y, big, "zoom \xC3\x97" + fmt_int(f.zoom), WHITE, 1);
            y += ch * big * 1.05f;
            h.text(xr, y, small, "field of view " + fmt_uas(width_m * uas_m), ACCENT, 1);
            y += lh;
            h.text(xr, y, small, "= " + fmt_g(width_m, "%.3g") + " M (gravitational radii)", DIM, 1);
            y += lh;
        } else {
            h.text(xr, y, big, fmt_g(S.nu_obs / 1e9, "%.0f") + " GHz", WHITE, 1);
            y += ch * big * 1.05f;
            const double beam = c.num("post.beam_uas", 20.0) * f.blur;
            h.text(xr, y, small,
                   beam > 0.5 ? "resolution " + fmt_g(beam, "%.0f") + " \xC2\xB5" "as (Event Horizon Telescope)"
                              : "resolution: unlimited",
                   ACCENT, 1);
            y += lh;
        }
        const double days = f.time_m * si.tg_s / 86400.0;
        h.text(xr, y, small, "elapsed " + fmt_g(days, days < 10 ? "%.1f" : "%.0f") + " days", DIM, 1);
    } else {
        const double km = f.r * si.rg_cm * 1e-5;
        h.text(xr, y, big, "r = " + fmt_g(f.r, f.r < 100 ? "%.1f" : "%.0f") + " M", WHITE, 1);
        y += ch * big * 1.05f;
        const std::string dist = km < 1e8 ? fmt_int(km) + " km"
                                          : fmt_g(km / 1.495978707e8, km < 1.5e10 ? "%.1f" : "%.0f") + " AU";
        h.text(xr, y, small, dist + " from the centre", ACCENT, 1);
        y += lh;
        const double v = std::sqrt(F.beta[0] * F.beta[0] + F.beta[1] * F.beta[1] + F.beta[2] * F.beta[2]);
        if (v > 1e-3) {
            h.text(xr, y, small, "orbital speed " + fmt_g(v, "%.2f") + " c", DIM, 1);
            y += lh;
        }
        h.text(xr, y, small, "inclination " + fmt_g(f.incl * 180 / 3.14159265358979323846, "%.0f") + "\xC2\xB0", DIM, 1);
    }
    h.end_block(pad);

    // bottom left: notes (replaced by the hidden-direct-image note)
    std::vector<std::string> notes = fields(c.str("hud.note"));
    if (f.hide_direct > 0.5) notes = {"direct image removed: only light that circled the black hole is shown"};
    if (!notes.empty() && !notes[0].empty()) {
        h.begin_block();
        float yb = (float)H - m - lh * (float)notes.size();
        for (const std::string& l : notes) {
            h.text(m, yb, small, l, DIM);
            yb += lh;
        }
        h.end_block(pad);
    }
    // bottom centre: photon-ring order legend while the order tint is on
    if (f.overlay > 0.05) {
        h.begin_block();
        const float yb = (float)H - m - lh * 2.2f;
        const float a0 = h.alpha_;
        h.alpha_ = a0 * (float)std::min(1.0, f.overlay * 1.5);
        h.text(0.5f * (float)W, yb - lh, small, "photon-ring order n: half-orbits the light made around the hole", DIM, 2);
        const float x = 0.5f * (float)W - 3 * 150 * s * 0.5f;
        for (int n = 1; n <= 3; ++n)
            h.text(x + (float)(n - 1) * 150 * s, yb + 0.2f * lh, mid, n < 3 ? "n = " + std::to_string(n) : "n = 3+",
                   ORDER_COLOUR[n]);
        h.end_block(pad);
        h.alpha_ = a0;
    }
    // bottom right: angular scale bar of a round length near a sixth of the image width
    if (width_m > 0 && uas_m > 0) {
        const double pix_uas = width_m * uas_m / W;
        const double target = pix_uas * W / 6.0, e = std::pow(10.0, std::floor(std::log10(target)));
        double nice = e;
        for (double k : {1.0, 2.0, 5.0, 10.0})
            if (k * e <= target) nice = k * e;
        const int bar = (int)std::lround(nice / pix_uas);
        const int x1 = (int)((float)W - m), x0 = x1 - bar, yy = (int)((float)H - m - lh * 0.4f);
        h.begin_block();
        h.rect(x0, yy, x1, yy + std::max(2, (int)(3 * s)), 0xE0FFFFFFu);
        h.rect(x0, yy - (int)(8 * s), x0 + std::max(2, (int)(2 * s)), yy + (int)(3 * s), 0xE0FFFFFFu);
        h.rect(x1 - std::max(2, (int)(2 * s)), yy - (int)(8 * s), x1, yy + (int)(3 * s), 0xE0FFFFFFu);
        h.text(0.5f * (float)(x0 + x1), (float)yy - lh * 1.2f, small, fmt_uas(nice), WHITE, 2);
        h.end_block(pad);
    }
    return h.out;
}

}  // namespace bh
