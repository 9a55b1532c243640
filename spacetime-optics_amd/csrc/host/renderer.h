/* renderer.h - C++ host interface mirroring bhrt's Renderer (blackholes/src/renderer.hpp:72-88 of the
   reference) on top of the C ABI (include/bhrt.h). Same call names and argument meaning; errors are
   reported through the return value and `err` instead of aborting. */
#pragma once
#include <string>
#include <vector>

#include "../../../include/bhrt.h"
#include "scene.h"
#include "shot.h"

namespace bh {

// HdrFrame (renderer.hpp:37-46): radio mode px = (I_nu, I_n1, I_n2, I_n>=3) per pixel
struct HdrFrame {
    int W = 0, H = 0;
    std::vector<float> px;
    double lum_percentile = 0, lum_max = 0;   // display statistics of the post step
};

class Renderer {
public:
    Renderer() = default;
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    bool init(int device, const bh_scene& scene, const SceneTables& tables, const bh_settings& rs, std::string* err);
    bool update_scene(const bh_scene& s);
    // Renderer::render (render.cu:755-770) without post: adaptive sampling when the settings allow
    // more than spp samples (spp_max > spp), else the plain mean of spp samples. opts selects
    // precision / emissivity / jitter.
    bool render(const bh_frame& f, const bh_trace_opts& opts, HdrFrame* out, bh_stats* st);
    // Renderer::render with its post step (render.cu:755-792): the trace above, then in radio mode
    // the beam blur (post.blur_sigma_px > 0.5) and the display statistics at `percentile`.
    bool render(const bh_frame& f, const bh_trace_opts& opts, const bh_post_params& post, double percentile,
                HdrFrame* out, bh_stats* st);
    // Renderer::tonemap (render.cu:811-846) of a radio frame to 8-bit RGB, without the HUD overlay.
    bool tonemap(const HdrFrame& in, const bh_tone_params& tp, std::vector<uint8_t>* rgb8);
    // ... with the HUD's rectangles and glyphs drawn in `font` on top (bh_tonemap_hud).
    bool tonemap(const HdrFrame& in, const bh_tone_params& tp, const std::vector<bh_hud_glyph>& glyphs,
                 const std::vector<bh_hud_rect>& rects, const bh_font* font, std::vector<uint8_t>* rgb8);
    // One persistent launch over several frames (each with its own scene): pixel-centre or sample-0
    // rays, images one after the other in `out`.
    bool render_batch(const std::vector<bh_batch_frame>& frames, const bh_trace_opts& opts,
                      std::vector<HdrFrame>* out, bh_stats* st);
    // Flux density (Jy) of an n x n, spp = 2 jittered quick render; < 0 on error.
    double radio_flux(bh_frame f, int n, double pix_sr);
    bool vacuum_fates(const std::vector<double>& photons, double tol, double r_escape, int max_steps,
                      std::vector<int>* fates);
    bh_ctx* ctx() const { return ctx_; }
    // hands the context to the caller (who then owns it and calls bh_destroy)
    bh_ctx* release() {
        bh_ctx* c = ctx_;
        ctx_ = nullptr;
        return c;
    }
    std::string error() const;

private:
    bh_ctx* ctx_ = nullptr;
    bh_scene S_{};
    bh_settings rs_{};
};

/* Per-frame exposure (main.cpp:304-319): key / exp(Gaussian-weighted mean of ln lum_percentile over
   frames within 3 sigma, frames with lp <= 0 skipped); 1 where no frame contributes. */
std::vector<double> smooth_exposure(const std::vector<double>& lp, double key, double sigma);

/* Flux normalisation (main.cpp:475-508): secant iteration on ln n0 until the quick render of
   frame f0 (zoom 1, no target/offset, fov = info.normalize_fov_m, 160^2, spp 2) gives
   info.radio_flux_target_jy within |d ln F| < 0.01. Updates S.n0 and the renderer's scene. */
bool normalize_flux(Renderer& R, const Shot& shot, ShotFrame f0, const SceneInfo& info, bh_scene* S,
                    double* got_jy, std::string* err);

/* Self-test (main.cpp:323-372): shadow-edge rows (a, incl, traced -, traced +, analytic -, analytic +)
   for a in {1e-6, 0.5, 0.94, 0.99} x incl in {90, 17} deg; the traced edges come from the device
   bisection of find_critical_points. */
bool selftest_rows(Renderer& R, std::vector<std::vector<double>>* rows, std::string* err);

}  // namespace bh
