/* post.cpp - CPU restatement of bhrt's radio post step and tone map. TEST INFRASTRUCTURE ONLY
   (the same rules as oracle.h: only tests/ load it, as the checker of the GPU kernels).

   Follows blackholes/src/render.cu operation for operation in float arithmetic, compiled with
   -ffp-contract=off:
   - or_post_blur: Renderer::Impl::radio_blur (render.cu:742-753) - pyramid of 2 x 2 box means
     (k_down, render.cu:244-255, levels as render.cu:638-647), separable Gaussian with edge clamp
     normalised by its weight sum (k_gauss, render.cu:297-310), bilinear upsample (k_upsample and
     bilinear, render.cu:257-270,312-316). Tap weights exp(-i^2 / 2 s^2) in double rounded to float
     (the reference evaluates __expf per tap).
   - or_frame_stats: the displayed-intensity statistics of Renderer::render (render.cu:771-791).
   - or_tonemap_radio: k_tonemap's radio branch with colormap, burn, order tint, vignette
     (smoothstep, common.hpp:47-50), grain and dithered rounding (render.cu:334-407); the display
     gamma through bh_math.h's FP64 exp/log (the reference calls powf).
   - or_hud_draw: the HUD rectangles and glyphs blended onto the 8-bit image (render.cu:409-446,
     811-846), every element in list order (the reference draws the elements of one kernel launch
     concurrently, so where two overlap its result depends on scheduling). */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../spacetime-optics_amd/csrc/bh_math.h"

namespace {

struct F4 {
    float x, y, z, w;
};

float sat(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
float lerp(float a, float b, float t) { return a + (b - a) * t; }

uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
float rnd(uint32_t& s) {
    s = pcg(s);
    return (float)(s >> 8) * (1.0f / 16777216.0f);
}

void down(const std::vector<F4>& src, int sw, int sh, std::vector<F4>& dst, int dw, int dh) {
    dst.assign((size_t)dw * dh, F4{0, 0, 0, 0});
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            F4 a{0.f, 0.f, 0.f, 0.f};
            for (int j = 0; j < 2; ++j)
                for (int i = 0; i < 2; ++i) {
                    const F4& v = src[(size_t)std::min(2 * y + j, sh - 1) * sw + std::min(2 * x + i, sw - 1)];
                    a.x += v.x;
                    a.y += v.y;
                    a.z += v.z;
                    a.w += v.w;
                }
            dst[(size_t)y * dw + x] = F4{0.25f * a.x, 0.25f * a.y, 0.25f * a.z, 0.25f * a.w};
        }
}

void gauss(const std::vector<F4>& src, std::vector<F4>& dst, int w, int h, const std::vector<float>& wt, int R,
           bool horiz) {
    dst.assign((size_t)w * h, F4{0, 0, 0, 0});
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            F4 a{0.f, 0.f, 0.f, 0.f};
            float ws = 0.f;
            for (int i = -R; i <= R; ++i) {
                const int xx = horiz ? std::min(std::max(x + i, 0), w - 1) : x;
                const int yy = horiz ? y : std::min(std::max(y + i, 0), h - 1);
                const float g = wt[(size_t)(i + R)];
                const F4& v = src[(size_t)yy * w + xx];
                a.x += g * v.x;
                a.y += g * v.y;
                a.z += g * v.z;
                a.w += g * v.w;
                ws += g;
            }
            dst[(size_t)y * w + x] = F4{a.x / ws, a.y / ws, a.z / ws, a.w / ws};
        }
}

F4 bilinear(const std::vector<F4>& L, int w, int h, float x, float y) {
    x = std::fmin(std::fmax(x - 0.5f, 0.f), (float)w - 1.001f);
    y = std::fmin(std::fmax(y - 0.5f, 0.f), (float)h - 1.001f);
    const int ix = (int)x, iy = (int)y;
    const float fx = x - (float)ix, fy = y - (float)iy;
    const int ix1 = std::min(ix + 1, w - 1), iy1 = std::min(iy + 1, h - 1);
    const F4 &a = L[(size_t)iy * w + ix], &b = L[(size_t)iy * w + ix1], &c = L[(size_t)iy1 * w + ix],
             &d = L[(size_t)iy1 * w + ix1];
    return F4{(1 - fy) * ((1 - fx) * a.x + fx * b.x) + fy * ((1 - fx) * c.x + fx * d.x),
              (1 - fy) * ((1 - fx) * a.y + fx * b.y) + fy * ((1 - fx) * c.y + fx * d.y),
              (1 - fy) * ((1 - fx) * a.z + fx * b.z) + fy * ((1 - fx) * c.z + fx * d.z),
              (1 - fy) * ((1 - fx) * a.w + fx * b.w) + fy * ((1 - fx) * c.w + fx * d.w)};
}

void colormap(int which, float x, float* c) {
    x = sat(x);
    if (which == 0) {
        c[0] = sat(2.f * x);
        c[1] = sat(2.f * x - 0.5f);
        c[2] = sat(2.f * x - 1.f);
        return;
    }
    static const float T[9][3] = {{0.0015f, 0.0005f, 0.014f}, {0.087f, 0.045f, 0.224f}, {0.258f, 0.039f, 0.406f},
                                  {0.416f, 0.090f, 0.433f},   {0.578f, 0.148f, 0.404f}, {0.735f, 0.216f, 0.330f},
                                  {0.865f, 0.317f, 0.226f},   {0.955f, 0.497f, 0.086f}, {0.988f, 0.998f, 0.645f}};
    float f = x * 8.f;
    const int i = std::min((int)f, 7);
    f -= (float)i;
    for (int k = 0; k < 3; ++k) c[k] = lerp(T[i][k], T[i + 1][k], f);
}

void order_tint(float n, float* t) {
    static const float C[5][3] = {{1.f, 1.f, 1.f}, {0.25f, 0.9f, 1.f}, {1.f, 0.3f, 0.95f}, {1.f, 0.82f, 0.2f},
                                  {0.55f, 1.f, 0.3f}};
    n = std::fmin(std::fmax(n, 0.f), 4.f);
    const int i = std::min((int)n, 3);
    const float f = n - (float)i;
    for (int k = 0; k < 3; ++k) t[k] = lerp(C[i][k], C[i + 1][k], f);
}

float smoothstep(float a, float b, float x) {
    const float t = sat((x - a) / (b - a));
    return t * t * (3.f - 2.f * t);
}

}  // namespace

extern "C" {

// In place on px (W*H*4 floats); sigma in pixels (no blur at or below 0.5).
void or_post_blur(float* px, int W, int H, double sigma) {
    if (!(sigma > 0.5)) return;
    const int MAXL = 9;
    std::vector<std::vector<F4>> lev;
    std::vector<int> lw, lh;
    for (int w = W, h = H; (int)lw.size() < MAXL && w >= 2 && h >= 2; w = (w + 1) / 2, h = (h + 1) / 2) {
        lw.push_back(w);
        lh.push_back(h);
    }
    if (lw.empty()) {
        lw.push_back(W);
        lh.push_back(H);
    }
    const int nlev = (int)lw.size();
    lev.resize((size_t)nlev);
    lev[0].resize((size_t)W * H);
    std::memcpy(lev[0].data(), px, sizeof(float) * 4 * (size_t)W * H);
    int k = 0;
    while (k + 1 < nlev && sigma / (double)(1 << (k + 1)) > 3.0) ++k;
    for (int j = 1; j <= k; ++j) down(lev[(size_t)j - 1], lw[(size_t)j - 1], lh[(size_t)j - 1], lev[(size_t)j],
                                      lw[(size_t)j], lh[(size_t)j]);
    const float s = (float)(sigma / (double)(1 << k));
    const int R = (int)std::ceil(3.0f * s);
    std::vector<float> wt(2 * (size_t)R + 1);
    for (int i = -R; i <= R; ++i) wt[(size_t)(i + R)] = (float)std::exp(-0.5 * (double)i * i / ((double)s * s));
    std::vector<F4> tmp;
    gauss(lev[(size_t)k], tmp, lw[(size_t)k], lh[(size_t)k], wt, R, true);
    gauss(tmp, lev[(size_t)k], lw[(size_t)k], lh[(size_t)k], wt, R, false);
    if (k > 0) {
        const std::vector<F4>& L = lev[(size_t)k];
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                lev[0][(size_t)y * W + x] = bilinear(L, lw[(size_t)k], lh[(size_t)k],
                                                     ((float)x + 0.5f) * (float)lw[(size_t)k] / (float)W,
                                                     ((float)y + 0.5f) * (float)lh[(size_t)k] / (float)H);
    }
    std::memcpy(px, lev[0].data(), sizeof(float) * 4 * (size_t)W * H);
}

// lum_percentile and lum_max of the displayed intensity (render.cu:771-791)
void or_frame_stats(const float* px, int W, int H, float hide_direct, double percentile, double* lp, double* lmax) {
    const size_t n = (size_t)W * H;
    std::vector<float> lum;
    double mx = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const float* c = px + 4 * i;
        const double Y = c[0] - hide_direct * std::max(c[0] - c[1] - c[2] - c[3], 0.f);
        if (Y > 0) {
            lum.push_back((float)Y);
            mx = std::max(mx, Y);
        }
    }
    std::vector<float> sig;
    for (float v : lum)
        if (v > mx * 1e-5) sig.push_back(v);
    double pv = 0.0;
    if (!sig.empty()) {
        const size_t idx = std::min(sig.size() - 1, (size_t)(percentile * (double)(sig.size() - 1)));
        std::nth_element(sig.begin(), sig.begin() + (long)idx, sig.end());
        pv = sig[idx];
    }
    *lp = pv;
    *lmax = mx;
}

// rgb: W*H*3 bytes. tp fields as bh_tone_params: colormap, exposure, gamma, overlay, hide_direct,
// fade, vignette, grain, seed.
void or_tonemap_radio(const float* px, int W, int H, int cmap, float exposure, float gamma, float overlay,
                      float hide_direct, float fade, float vignette, float grain, uint32_t seed, uint8_t* rgb) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int p = y * W + x;
            const float* v = px + 4 * (size_t)p;
            const float I0 = std::fmax(v[0] - v[1] - v[2] - v[3], 0.f);
            const float Id = std::fmax(v[0] - hide_direct * I0, 0.f);
            const float I = Id * exposure;
            float xx = I;
            if (gamma != 1.f) xx = I > 0.f ? (float)bhm::exp((double)gamma * bhm::log((double)I)) : 0.f;
            float c[3];
            colormap(cmap, xx, c);
            const float over = std::fmax(xx - 1.f, 0.f);
            for (int k = 0; k < 3; ++k) c[k] = std::fmin(c[k] + 0.25f * over, 1.f);
            if (overlay > 0.f && Id > 0.f) {
                const float f1 = sat(v[1] / Id), f2 = sat(v[2] / Id), f3 = sat(v[3] / Id), f0 = sat(1.f - f1 - f2 - f3);
                const float lum = std::fmax(0.3f * c[0] + 0.59f * c[1] + 0.11f * c[2], 0.f);
                float t1[3], t2[3], t3[3];
                order_tint(1.f, t1);
                order_tint(2.f, t2);
                order_tint(3.f, t3);
                const float L = sat(lum * 1.3f + 0.05f);
                float m[3];
                for (int k = 0; k < 3; ++k) m[k] = f0 * c[k] + L * (f1 * t1[k] + f2 * t2[k] + f3 * t3[k]);
                for (int k = 0; k < 3; ++k) c[k] = lerp(c[k], sat(m[k]), overlay);
            }
            if (vignette > 0.f) {
                const float dx = ((float)x + 0.5f) / (float)W - 0.5f;
                const float dy = (((float)y + 0.5f) / (float)H - 0.5f) * (float)H / (float)W;
                const float v2 = 1.f - vignette * smoothstep(0.25f, 0.75f, std::sqrt(dx * dx + dy * dy));
                for (int k = 0; k < 3; ++k) c[k] *= v2;
            }
            uint32_t hs = pcg((uint32_t)p ^ seed);
            const float u0 = rnd(hs), u1 = rnd(hs);
            const float gr = grain * (u0 + u1 - 1.f);
            float d[3];
            for (int k = 0; k < 3; ++k) {
                const float a = rnd(hs), b = rnd(hs);
                d[k] = a - b;
            }
            for (int k = 0; k < 3; ++k) {
                const float q = (c[k] + gr) * fade;
                rgb[3 * (size_t)p + k] = (uint8_t)std::fmin(std::fmax(q * 255.f + 0.5f + d[k], 0.f), 255.f);
            }
        }
}


// HUD on an 8-bit RGB image (Renderer::tonemap, render.cu:811-846, with k_hud_rects and
// k_hud_glyphs, render.cu:409-446): the rectangles stable-sorted by layer, then every glyph's shadow
// (offset off px, black, 3/4 of the glyph's alpha), then every glyph, each blended in list order.
// glyphs: ng records {float x, y, scale; int glyph; uint32 rgba}; rects: nr records {int x0, y0, x1,
// y1; uint32 rgba; int layer}; atlas: glyphs x ch x cw bytes.
struct OrGlyph {
    float x, y, scale;
    int glyph;
    uint32_t rgba;
};
struct OrRect {
    int x0, y0, x1, y1;
    uint32_t rgba;
    int layer;
};

void or_hud_draw(uint8_t* rgb, int W, int H, const OrGlyph* glyphs, int ng, const OrRect* rects_in, int nr,
                 const uint8_t* atlas, int cw_cell, int ch_cell, float off) {
    std::vector<OrRect> rects(rects_in, rects_in + nr);
    std::stable_sort(rects.begin(), rects.end(), [](const OrRect& a, const OrRect& b) { return a.layer < b.layer; });
    for (const OrRect& R : rects) {
        const float a = (float)((R.rgba >> 24) & 255) / 255.f;
        const int s[3] = {(int)(R.rgba & 255), (int)((R.rgba >> 8) & 255), (int)((R.rgba >> 16) & 255)};
        for (int Y = std::max(R.y0, 0); Y < std::min(R.y1, H); ++Y)
            for (int X = std::max(R.x0, 0); X < std::min(R.x1, W); ++X) {
                uint8_t* o = rgb + 3 * ((size_t)Y * W + X);
                for (int k = 0; k < 3; ++k) {
                    const float d = (float)o[k];
                    o[k] = (uint8_t)std::fmin(d * (1.f - a) + (float)s[k] * a + 0.5f, 255.f);
                }
            }
    }
    for (int pass = 0; pass < 2; ++pass) {
        const bool shadow = pass == 0;
        const float dx = shadow ? off : 0.f, dy = shadow ? off : 0.f;
        for (int gi = 0; gi < ng; ++gi) {
            const OrGlyph& G = glyphs[gi];
            const int cw = (int)std::ceil((float)cw_cell * G.scale) + 1, ch = (int)std::ceil((float)ch_cell * G.scale) + 1;
            float c[3] = {(float)(G.rgba & 255) / 255.f, (float)((G.rgba >> 8) & 255) / 255.f,
                          (float)((G.rgba >> 16) & 255) / 255.f};
            float am = (float)((G.rgba >> 24) & 255) / 255.f;
            if (shadow) {
                c[0] = c[1] = c[2] = 0.f;
                am *= 0.75f;
            }
            const int ox = (int)std::floor(G.x + dx), oy = (int)std::floor(G.y + dy);
            auto A = [&](int xx, int yy) -> float {
                if (xx < 0 || yy < 0 || xx >= cw_cell || yy >= ch_cell) return 0.f;
                return (float)atlas[((size_t)G.glyph * ch_cell + yy) * cw_cell + xx] / 255.f;
            };
            for (int iy = 0; iy < ch; ++iy)
                for (int ix = 0; ix < cw; ++ix) {
                    const int X = ox + ix, Y = oy + iy;
                    if (X < 0 || Y < 0 || X >= W || Y >= H) continue;
                    const float u = ((float)ix + 0.5f) / G.scale - 0.5f, v = ((float)iy + 0.5f) / G.scale - 0.5f;
                    const int u0 = (int)std::floor(u), v0 = (int)std::floor(v);
                    const float fu = u - (float)u0, fv = v - (float)v0;
                    const float a = ((1.f - fv) * ((1.f - fu) * A(u0, v0) + fu * A(u0 + 1, v0)) +
                                     fv * ((1.f - fu) * A(u0, v0 + 1) + fu * A(u0 + 1, v0 + 1))) * am;
                    if (a <= 0.f) continue;
                    uint8_t* o = rgb + 3 * ((size_t)Y * W + X);
                    for (int k = 0; k < 3; ++k) {
                        const float d = (float)o[k];
                        o[k] = (uint8_t)std::fmin(d * (1.f - a) + c[k] * 255.f * a + 0.5f, 255.f);
                    }
                }
        }
    }
}

}  // extern "C"
