/* post.hip - radio post-processing and display on gfx950: the interferometer-beam blur
   (Renderer::Impl::radio_blur, blackholes/src/render.cu:742-753, with the pyramid filters at
   render.cu:244-330), the frame's display statistics (render.cu:771-791) and the radio tone map
   (k_tonemap, render.cu:360-407, radio branch). C ABI: bh_post_radio, bh_tonemap (include/bhrt.h).

   Layout: frames are float4 pixels, row-major from the top left. Every kernel is one work-item per
   output pixel over a 1-D grid (coalesced rows); the pyramid levels and the tone-mapped bytes live
   in context buffers. All of it is memory-bound and small next to the trace (a 1024^2 frame is
   16 MB per pass).

   Arithmetic: the float operations of the reference in the same order, compiled with
   -ffp-contract=off, so the CPU restatement in the oracle reproduces every byte. Two deliberate
   differences: the Gaussian tap weights exp(-i^2 / 2 sigma^2) are computed once on the host in
   double and rounded to float (the reference evaluates the approximate __expf per tap), and the
   display gamma uses the FP64 exp/log of bh_math.h (the reference calls powf; gamma = 1, the
   default, is exact either way). */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "bh_math.h"
#include "ctx.h"

namespace {

#define PCK(ctx, x)                                                                                        \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) {                                                                            \
            (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);                                   \
            return BH_ERR_HIP;                                                                             \
        }                                                                                                  \
    } while (0)

constexpr int POST_BLOCK = 256;
constexpr int POST_MAX_LEVELS = 9;   // pyramid depth of the reference (render.cu:32)

__device__ __forceinline__ float sat(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
__device__ __forceinline__ float mix(float a, float b, float t) { return a + (b - a) * t; }

// Next pyramid level: the 2 x 2 box mean, the last row/column repeated at odd sizes.
__global__ __launch_bounds__(POST_BLOCK) void k_pyr_down(const float4* __restrict__ src, int sw, int sh,
                                                         float4* __restrict__ dst, int dw, int dh) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= dw * dh) return;
    const int x = p % dw, y = p / dw;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = 0; j < 2; ++j) {
        const int yy = min(2 * y + j, sh - 1);
        for (int i = 0; i < 2; ++i) {
            const float4 v = src[yy * sw + min(2 * x + i, sw - 1)];
            a.x += v.x;
            a.y += v.y;
            a.z += v.z;
            a.w += v.w;
        }
    }
    dst[p] = make_float4(0.25f * a.x, 0.25f * a.y, 0.25f * a.z, 0.25f * a.w);
}

// One pass of the separable Gaussian (taps -R..R, clamped at the edges, normalised by the weight sum).
__global__ __launch_bounds__(POST_BLOCK) void k_gauss_pass(const float4* __restrict__ src, float4* __restrict__ dst,
                                                           int w, int h, const float* __restrict__ wt, int R,
                                                           int horizontal) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= w * h) return;
    const int x = p % w, y = p / w;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    float ws = 0.f;
    for (int i = -R; i <= R; ++i) {
        const int xx = horizontal ? min(max(x + i, 0), w - 1) : x;
        const int yy = horizontal ? y : min(max(y + i, 0), h - 1);
        const float g = wt[i + R];
        const float4 v = src[yy * w + xx];
        a.x += g * v.x;
        a.y += g * v.y;
        a.z += g * v.z;
        a.w += g * v.w;
        ws += g;
    }
    dst[p] = make_float4(a.x / ws, a.y / ws, a.z / ws, a.w / ws);
}

// Bilinear read of level L at continuous pixel coordinates (pixel centres at +0.5), edge-clamped.
__device__ __forceinline__ float4 bilinear(const float4* __restrict__ L, int w, int h, float x, float y) {
    x = fminf(fmaxf(x - 0.5f, 0.f), (float)w - 1.001f);
    y = fminf(fmaxf(y - 0.5f, 0.f), (float)h - 1.001f);
    const int ix = (int)x, iy = (int)y;
    const float fx = x - (float)ix, fy = y - (float)iy;
    const int ix1 = min(ix + 1, w - 1), iy1 = min(iy + 1, h - 1);
    const float4 a = L[iy * w + ix], b = L[iy * w + ix1], c = L[iy1 * w + ix], d = L[iy1 * w + ix1];
    const float gx = 1.f - fx, gy = 1.f - fy;
    return make_float4(gy * (gx * a.x + fx * b.x) + fy * (gx * c.x + fx * d.x),
                       gy * (gx * a.y + fx * b.y) + fy * (gx * c.y + fx * d.y),
                       gy * (gx * a.z + fx * b.z) + fy * (gx * c.z + fx * d.z),
                       gy * (gx * a.w + fx * b.w) + fy * (gx * c.w + fx * d.w));
}

__global__ __launch_bounds__(POST_BLOCK) void k_pyr_up(const float4* __restrict__ L, int lw, int lh,
                                                       float4* __restrict__ img, int W, int H) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W * H) return;
    const int x = p % W, y = p / W;
    img[p] = bilinear(L, lw, lh, ((float)x + 0.5f) * (float)lw / (float)W, ((float)y + 0.5f) * (float)lh / (float)H);
}

// PCG hash and uniform float of render.cu / device_util.cuh:15-23
__device__ __forceinline__ unsigned int pcg_hash(unsigned int v) {
    const unsigned int s = v * 747796405u + 2891336453u;
    const unsigned int w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float rnd01(unsigned int& s) {
    s = pcg_hash(s);
    return (float)(s >> 8) * (1.0f / 16777216.0f);
}

// Radio colormaps on [0, 1]: afmhot (the EHT palette) and an 8-segment inferno (render.cu:334-345)
__device__ __forceinline__ float3 colormap(int which, float x) {
    x = sat(x);
    if (which == 0) return make_float3(sat(2.f * x), sat(2.f * x - 0.5f), sat(2.f * x - 1.f));
    const float T[9][3] = {{0.0015f, 0.0005f, 0.014f}, {0.087f, 0.045f, 0.224f}, {0.258f, 0.039f, 0.406f},
                           {0.416f, 0.090f, 0.433f},   {0.578f, 0.148f, 0.404f}, {0.735f, 0.216f, 0.330f},
                           {0.865f, 0.317f, 0.226f},   {0.955f, 0.497f, 0.086f}, {0.988f, 0.998f, 0.645f}};
    float f = x * 8.f;
    const int i = min((int)f, 7);
    f -= (float)i;
    return make_float3(mix(T[i][0], T[i + 1][0], f), mix(T[i][1], T[i + 1][1], f), mix(T[i][2], T[i + 1][2], f));
}

// Photon-ring order tints n = 1, 2, 3 (render.cu:348-355 at integer n)
__constant__ float c_tint[3][3] = {{0.25f, 0.9f, 1.f}, {1.f, 0.3f, 0.95f}, {1.f, 0.82f, 0.2f}};

__global__ __launch_bounds__(POST_BLOCK) void k_tone_radio(const float4* __restrict__ hdr, int W, int H,
                                                           bh_tone_params tp, unsigned char* __restrict__ rgb) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= W * H) return;
    const int x = p % W, y = p / W;
    const float4 v = hdr[p];
    const float I0 = fmaxf(v.x - v.y - v.z - v.w, 0.f);
    const float Id = fmaxf(v.x - tp.hide_direct * I0, 0.f);
    const float I = Id * tp.exposure;
    float xx = I;
    if (tp.gamma != 1.f) xx = I > 0.f ? (float)bhm::exp((double)tp.gamma * bhm::log((double)I)) : 0.f;
    float3 c = colormap(tp.colormap, xx);
    const float over = fmaxf(xx - 1.f, 0.f);   // highlights burn to white
    c = make_float3(fminf(c.x + 0.25f * over, 1.f), fminf(c.y + 0.25f * over, 1.f), fminf(c.z + 0.25f * over, 1.f));
    if (tp.overlay > 0.f && Id > 0.f) {
        const float f1 = sat(v.y / Id), f2 = sat(v.z / Id), f3 = sat(v.w / Id), f0 = sat(1.f - f1 - f2 - f3);
        const float lum = fmaxf(0.3f * c.x + 0.59f * c.y + 0.11f * c.z, 0.f);
        const float L = sat(lum * 1.3f + 0.05f);
        const float mx = f0 * c.x + L * (f1 * c_tint[0][0] + f2 * c_tint[1][0] + f3 * c_tint[2][0]);
        const float my = f0 * c.y + L * (f1 * c_tint[0][1] + f2 * c_tint[1][1] + f3 * c_tint[2][1]);
        const float mz = f0 * c.z + L * (f1 * c_tint[0][2] + f2 * c_tint[1][2] + f3 * c_tint[2][2]);
        const float k = tp.overlay;
        c = make_float3(mix(c.x, sat(mx), k), mix(c.y, sat(my), k), mix(c.z, sat(mz), k));
    }
    if (tp.vignette > 0.f) {
        const float dx = ((float)x + 0.5f) / (float)W - 0.5f, dy = (((float)y + 0.5f) / (float)H - 0.5f) * (float)H / (float)W;
        const float t = sat((__builtin_sqrtf(dx * dx + dy * dy) - 0.25f) / (0.75f - 0.25f));
        const float v2 = 1.f - tp.vignette * (t * t * (3.f - 2.f * t));
        c.x *= v2;
        c.y *= v2;
        c.z *= v2;
    }
    unsigned int hs = pcg_hash((unsigned int)p ^ tp.seed);
    const float u0 = rnd01(hs), u1 = rnd01(hs);
    const float gr = tp.grain * (u0 + u1 - 1.f);
    float d[3];
    for (int k = 0; k < 3; ++k) {
        const float a = rnd01(hs), b = rnd01(hs);
        d[k] = a - b;
    }
    const float cc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; ++k) {
        const float q = (cc[k] + gr) * tp.fade;
        rgb[3 * (size_t)p + k] = (unsigned char)fminf(fmaxf(q * 255.f + 0.5f + d[k], 0.f), 255.f);
    }
}

int grow(bh_ctx* c, void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes && *p) return BH_OK;
    if (*p) {
        PCK(c, hipStreamSynchronize(c->st));
        PCK(c, hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    PCK(c, hipMalloc(p, std::max<size_t>(bytes, 1)));
    *cap = bytes;
    return BH_OK;
}

inline dim3 grid1(long long n) { return dim3((unsigned int)((n + POST_BLOCK - 1) / POST_BLOCK)); }

// HUD over the tone-mapped bytes, one work-item per pixel: every rectangle (sorted by layer on the
// host), then every glyph's shadow, then every glyph, in list order - the blends of k_hud_rects and
// k_hud_glyphs (render.cu:409-446) in the same float operations. The element loops are wave-uniform,
// so each element's fields arrive through scalar loads; a glyph touches only the pixels of its cell.
__global__ __launch_bounds__(POST_BLOCK) void k_hud(unsigned char* __restrict__ rgb, int W, int H,
                                                    const bh_hud_rect* __restrict__ rects, int nr,
                                                    const bh_hud_glyph* __restrict__ glyphs, int ng,
                                                    const unsigned char* __restrict__ atlas, int cell_w, int cell_h,
                                                    float off) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)W * H) return;
    const int X = (int)(p % W), Y = (int)(p / W);
    unsigned char* o = rgb + 3 * p;
    float d0 = (float)o[0], d1 = (float)o[1], d2 = (float)o[2];
    bool touched = false;
    for (int i = 0; i < nr; ++i) {
        const bh_hud_rect R = rects[i];
        if (X < R.x0 || X >= R.x1 || Y < R.y0 || Y >= R.y1) continue;
        const float a = (float)((R.rgba >> 24) & 255u) / 255.f;
        d0 = (float)(unsigned char)fminf(d0 * (1.f - a) + (float)(int)(R.rgba & 255u) * a + 0.5f, 255.f);
        d1 = (float)(unsigned char)fminf(d1 * (1.f - a) + (float)(int)((R.rgba >> 8) & 255u) * a + 0.5f, 255.f);
        d2 = (float)(unsigned char)fminf(d2 * (1.f - a) + (float)(int)((R.rgba >> 16) & 255u) * a + 0.5f, 255.f);
        touched = true;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const float dxy = pass == 0 ? off : 0.f;
        for (int i = 0; i < ng; ++i) {
            const bh_hud_glyph G = glyphs[i];
            const int cw = (int)ceilf((float)cell_w * G.scale) + 1, ch = (int)ceilf((float)cell_h * G.scale) + 1;
            const int ix = X - (int)floorf(G.x + dxy), iy = Y - (int)floorf(G.y + dxy);
            if (ix < 0 || iy < 0 || ix >= cw || iy >= ch) continue;
            float r = (float)(G.rgba & 255u) / 255.f, g = (float)((G.rgba >> 8) & 255u) / 255.f,
                  b = (float)((G.rgba >> 16) & 255u) / 255.f, am = (float)((G.rgba >> 24) & 255u) / 255.f;
            if (pass == 0) {
                r = g = b = 0.f;
                am *= 0.75f;
            }
            const float u = ((float)ix + 0.5f) / G.scale - 0.5f, v = ((float)iy + 0.5f) / G.scale - 0.5f;
            const int u0 = (int)floorf(u), v0 = (int)floorf(v);
            const float fu = u - (float)u0, fv = v - (float)v0;
            const unsigned char* cell = atlas + (size_t)G.glyph * cell_h * cell_w;
            auto A = [&](int xx, int yy) -> float {
                return (xx < 0 || yy < 0 || xx >= cell_w || yy >= cell_h) ? 0.f : (float)cell[yy * cell_w + xx] / 255.f;
            };
            const float a = ((1.f - fv) * ((1.f - fu) * A(u0, v0) + fu * A(u0 + 1, v0)) +
                             fv * ((1.f - fu) * A(u0, v0 + 1) + fu * A(u0 + 1, v0 + 1))) * am;
            if (a <= 0.f) continue;
            d0 = (float)(unsigned char)fminf(d0 * (1.f - a) + r * 255.f * a + 0.5f, 255.f);
            d1 = (float)(unsigned char)fminf(d1 * (1.f - a) + g * 255.f * a + 0.5f, 255.f);
            d2 = (float)(unsigned char)fminf(d2 * (1.f - a) + b * 255.f * a + 0.5f, 255.f);
            touched = true;
        }
    }
    if (touched) {
        o[0] = (unsigned char)d0;
        o[1] = (unsigned char)d1;
        o[2] = (unsigned char)d2;
    }
}

}  // namespace

extern "C" {

int bh_post_radio(bh_ctx* c, float* px, int W, int H, const bh_post_params* params, double percentile,
                  double* lum_percentile, double* lum_max) {
    if (!c || !px || W <= 0 || H <= 0) return BH_ERR_ARG;
    if ((long long)W * H > (1ll << 28)) {
        c->err = "frame too large for the post step (2^28 pixels)";
        return BH_ERR_ARG;
    }
    PCK(c, hipSetDevice(c->device));
    const bh_post_params pp = params ? *params : bh_post_params{0.0, 0.f};
    const size_t npix = (size_t)W * H;
    if (pp.blur_sigma_px > 0.5) {
        // pyramid of the reference: halve (rounding up) while both sides are >= 2, at most 9 levels
        int lw[POST_MAX_LEVELS], lh[POST_MAX_LEVELS], nlev = 0;
        size_t loff[POST_MAX_LEVELS], total = 0;
        for (int w = W, h = H; nlev < POST_MAX_LEVELS && w >= 2 && h >= 2; ++nlev, w = (w + 1) / 2, h = (h + 1) / 2) {
            lw[nlev] = w;
            lh[nlev] = h;
            loff[nlev] = total;
            total += (size_t)w * h;
        }
        if (nlev == 0) {   // a 1-pixel-wide frame: the reference has no level to blur on
            lw[0] = W;
            lh[0] = H;
            loff[0] = 0;
            total = npix;
            nlev = 1;
        }
        // the level the blur runs on: the coarsest whose sigma stays above 3 pixels (render.cu:744)
        int k = 0;
        while (k + 1 < nlev && pp.blur_sigma_px / (double)(1 << (k + 1)) > 3.0) ++k;
        const float s = (float)(pp.blur_sigma_px / (double)(1 << k));
        const int R = (int)std::ceil(3.0f * s);
        std::vector<float> wt(2 * (size_t)R + 1);
        for (int i = -R; i <= R; ++i) wt[(size_t)(i + R)] = (float)std::exp(-0.5 * (double)i * i / ((double)s * s));
        int rc;
        if ((rc = grow(c, (void**)&c->d_post_lev, &c->post_lev_cap, sizeof(float4) * total)) ||
            (rc = grow(c, (void**)&c->d_post_tmp, &c->post_tmp_cap, sizeof(float4) * (size_t)lw[k] * lh[k])) ||
            (rc = grow(c, (void**)&c->d_post_wt, &c->post_wt_cap, sizeof(float) * wt.size())))
            return rc;
        float4* lev = (float4*)c->d_post_lev;
        PCK(c, hipMemcpyAsync(lev, px, sizeof(float4) * npix, hipMemcpyHostToDevice, c->st));
        PCK(c, hipMemcpyAsync(c->d_post_wt, wt.data(), sizeof(float) * wt.size(), hipMemcpyHostToDevice, c->st));
        for (int j = 1; j <= k; ++j)
            hipLaunchKernelGGL(k_pyr_down, grid1((long long)lw[j] * lh[j]), dim3(POST_BLOCK), 0, c->st,
                               lev + loff[j - 1], lw[j - 1], lh[j - 1], lev + loff[j], lw[j], lh[j]);
        float4* L = lev + loff[k];
        float4* tmp = (float4*)c->d_post_tmp;
        const long long nk = (long long)lw[k] * lh[k];
        hipLaunchKernelGGL(k_gauss_pass, grid1(nk), dim3(POST_BLOCK), 0, c->st, L, tmp, lw[k], lh[k],
                           (const float*)c->d_post_wt, R, 1);
        hipLaunchKernelGGL(k_gauss_pass, grid1(nk), dim3(POST_BLOCK), 0, c->st, tmp, L, lw[k], lh[k],
                           (const float*)c->d_post_wt, R, 0);
        if (k > 0) hipLaunchKernelGGL(k_pyr_up, grid1((long long)npix), dim3(POST_BLOCK), 0, c->st, L, lw[k], lh[k],
                                      lev, W, H);
        PCK(c, hipGetLastError());
        PCK(c, hipMemcpyAsync(px, lev, sizeof(float4) * npix, hipMemcpyDeviceToHost, c->st));
        PCK(c, hipStreamSynchronize(c->st));
    }
    // statistics of the displayed intensity (render.cu:771-791)
    if (lum_percentile || lum_max) {
        std::vector<float> lum;
        lum.reserve(npix);
        double mx = 0.0;
        for (size_t i = 0; i < npix; ++i) {
            const float* v = px + 4 * i;
            const double Y = v[0] - pp.hide_direct * std::max(v[0] - v[1] - v[2] - v[3], 0.f);
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
            const double q = std::min(std::max(percentile, 0.0), 1.0);
            const size_t idx = std::min(sig.size() - 1, (size_t)(q * (double)(sig.size() - 1)));
            std::nth_element(sig.begin(), sig.begin() + (long)idx, sig.end());
            pv = sig[idx];
        }
        if (lum_percentile) *lum_percentile = pv;
        if (lum_max) *lum_max = mx;
    }
    return BH_OK;
}

static int copy_rgb(bh_ctx* c, int W, int H, unsigned char* rgb8) {
    PCK(c, hipMemcpyAsync(rgb8, c->d_post_rgb, 3 * (size_t)W * H, hipMemcpyDeviceToHost, c->st));
    PCK(c, hipStreamSynchronize(c->st));
    return BH_OK;
}

// the radio tone map into the context's device byte buffer
static int tone_device(bh_ctx* c, const float* px, int W, int H, const bh_tone_params* params) {
    if ((long long)W * H > (1ll << 28)) {
        c->err = "frame too large for the tone map (2^28 pixels)";
        return BH_ERR_ARG;
    }
    if (params->colormap != 0 && params->colormap != 1) {
        c->err = "colormap must be 0 (afmhot) or 1 (inferno)";
        return BH_ERR_ARG;
    }
    PCK(c, hipSetDevice(c->device));
    const size_t npix = (size_t)W * H;
    int rc;
    if ((rc = grow(c, (void**)&c->d_post_tmp, &c->post_tmp_cap, sizeof(float4) * npix)) ||
        (rc = grow(c, (void**)&c->d_post_rgb, &c->post_rgb_cap, 3 * npix)))
        return rc;
    PCK(c, hipMemcpyAsync(c->d_post_tmp, px, sizeof(float4) * npix, hipMemcpyHostToDevice, c->st));
    hipLaunchKernelGGL(k_tone_radio, grid1((long long)npix), dim3(POST_BLOCK), 0, c->st,
                       (const float4*)c->d_post_tmp, W, H, *params, (unsigned char*)c->d_post_rgb);
    PCK(c, hipGetLastError());
    return BH_OK;
}

int bh_tonemap(bh_ctx* c, const float* px, int W, int H, const bh_tone_params* params, unsigned char* rgb8) {
    if (!c || !px || !rgb8 || !params || W <= 0 || H <= 0) return BH_ERR_ARG;
    const int rc = tone_device(c, px, W, H, params);
    return rc ? rc : copy_rgb(c, W, H, rgb8);
}

int bh_tonemap_hud(bh_ctx* c, const float* px, int W, int H, const bh_tone_params* params, const bh_hud_glyph* glyphs,
                   int n_glyphs, const bh_hud_rect* rects, int n_rects, const bh_font* font, unsigned char* rgb8) {
    if (!c || !px || !rgb8 || !params || W <= 0 || H <= 0 || n_glyphs < 0 || n_rects < 0) return BH_ERR_ARG;
    if ((n_glyphs > 0 && !glyphs) || (n_rects > 0 && !rects)) return BH_ERR_ARG;
    const int ng = std::min(n_glyphs, 4096), nr = std::min(n_rects, 1024);   // the reference's limits
    if (ng > 0) {
        if (!font || !font->atlas || font->cell_w < 1 || font->cell_h < 1 || font->glyphs < 1 ||
            font->cell_w > 256 || font->cell_h > 256) {
            c->err = "HUD glyphs need a font (atlas, cell_w and cell_h in [1, 256], glyphs >= 1)";
            return BH_ERR_ARG;
        }
        for (int i = 0; i < ng; ++i) {
            if (glyphs[i].glyph < 0 || glyphs[i].glyph >= font->glyphs) {
                c->err = "HUD glyph index outside the font";
                return BH_ERR_ARG;
            }
            if (!(glyphs[i].scale > 0.f && glyphs[i].scale <= 64.f) || !std::isfinite(glyphs[i].x) ||
                !std::isfinite(glyphs[i].y)) {
                c->err = "HUD glyph scale must be in (0, 64] and its position finite";
                return BH_ERR_ARG;
            }
        }
    }
    int rc = tone_device(c, px, W, H, params);
    if (rc) return rc;
    if (ng == 0 && nr == 0) return copy_rgb(c, W, H, rgb8);
    std::vector<bh_hud_rect> sorted(rects, rects + nr);
    std::stable_sort(sorted.begin(), sorted.end(), [](const bh_hud_rect& a, const bh_hud_rect& b) { return a.layer < b.layer; });
    const size_t rbytes = sizeof(bh_hud_rect) * (size_t)nr, gbytes = sizeof(bh_hud_glyph) * (size_t)ng;
    const size_t abytes = ng > 0 ? (size_t)font->glyphs * font->cell_h * font->cell_w : 0;
    const size_t goff = (rbytes + 255) & ~(size_t)255, aoff = goff + ((gbytes + 255) & ~(size_t)255);
    if ((rc = grow(c, &c->d_hud, &c->hud_cap, aoff + abytes + 1))) return rc;
    unsigned char* base = (unsigned char*)c->d_hud;
    if (nr) PCK(c, hipMemcpyAsync(base, sorted.data(), rbytes, hipMemcpyHostToDevice, c->st));
    if (ng) PCK(c, hipMemcpyAsync(base + goff, glyphs, gbytes, hipMemcpyHostToDevice, c->st));
    if (ng) PCK(c, hipMemcpyAsync(base + aoff, font->atlas, abytes, hipMemcpyHostToDevice, c->st));
    const float off = std::max(1.5f, (float)H / 720.f);
    hipLaunchKernelGGL(k_hud, grid1((long long)W * H), dim3(POST_BLOCK), 0, c->st, (unsigned char*)c->d_post_rgb, W, H,
                       (const bh_hud_rect*)base, nr, (const bh_hud_glyph*)(base + goff), ng, base + aoff,
                       ng ? font->cell_w : 1, ng ? font->cell_h : 1, off);
    PCK(c, hipGetLastError());
    return copy_rgb(c, W, H, rgb8);
}

}  // extern "C"
