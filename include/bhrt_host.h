/* bhrt_host.h - C entry points of the standalone host layer (the bhrt CLI's building blocks): shot
   .ini parsing, scene and frame construction, flux normalisation and the self-test. An integrator
   that keeps the reference's own host code needs only bhrt.h; these exist for the bhrt executable
   and for tests that compare the host layer with the reference's.

   Reference counterparts (paths relative to the reference repository root):
     Config::load/str                 common/config.cpp:14-68
     build_scene                      blackholes/src/scene_build.cpp:83-214
     build_shot / make_frame          blackholes/src/shot.cpp:73-146
     find_critical_point              blackholes/src/shot.cpp:148-164
     flux normalisation               blackholes/src/main.cpp:475-508
     selftest                         blackholes/src/main.cpp:323-372 */
#ifndef BHRT_HOST_H
#define BHRT_HOST_H

#include "bhrt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Raw value of `key` from a shot .ini (includes resolved) plus "k=v;k=v" overrides.
   Returns 1 if found, 0 if absent, < 0 on a parse error (text in buf). */
int bhh_config_str(const char* ini, const char* sets, const char* key, char* buf, int len);

/* Scene scalars, the K2 table (k2 must hold 512 floats) and SceneInfo
   (info = mass_msun, distance_cm, rg_cm, uas_per_m, flux_target_jy, normalize_fov_m). */
int bhh_build_scene(const char* ini, const char* sets, bh_scene* scene, float* k2, float* k2_l0, float* k2_dl,
                    double* info, bh_trace_opts* opts, char* err, size_t errlen);

/* Turbulence grid of the [turbulence] section (see bh_tables); out holds n^3 * nt floats. */
void bhh_turbulence(int n, int nt, double half, double sigma, double corr_m, double corr_t, double dt,
                    uint32_t seed, float* out);

/* Frames of a shot: BHH_SHOT_VALUES values per frame (t, r, incl, phi, roll, yaw, pitch, fov,
   fov_in_m, zoom, target, off_x, off_y, time_m, blur, overlay, hide_direct, exposure_ev, fade, hud).
   Returns the frame count (< 0 on error). */
#define BHH_SHOT_VALUES 20
int bhh_build_shot(const char* ini, const char* sets, double* out, int maxn);

/* make_frame for explicit camera parameters (angles in radians). */
void bhh_make_frame(double a, double r, double incl, double phi, double roll, double yaw, double pitch, double fov,
                    int fov_in_m, double zoom, double target, double target_x, double target_y, double off_x,
                    double off_y, double time_m, int orbit, int W, int H, int spp, uint32_t seed, bh_frame* out);

/* Context for a shot .ini: builds the scene, tables and turbulence and creates the device context.
   With normalize != 0 and flow.flux_jy > 0 in radio mode, runs the flux normalisation first
   (*jy = the flux it reached). `frame` (nullable) receives frame `frame_index` of the shot at the
   camera.width x camera.height resolution (zoom target resolved). */
int bhh_create_from_ini(const char* ini, const char* sets, int device, int normalize, int frame_index,
                        bh_ctx** ctx, bh_scene* scene, bh_frame* frame, bh_trace_opts* opts, double* jy, char* err,
                        size_t errlen);

/* find_critical_point along psi for a static frame, on the context's spin. */
int bhh_find_critical_point(bh_ctx* ctx, const bh_frame* frame, double psi, double* rho);

/* 8-bit PNG of w x h pixels with `channels` (1, 3 or 4) bytes each, row-major from the top left,
   Sub-filtered scanlines, zlib `level` (write_png, common/image_io.cpp:37-66). Returns 0 or -1. */
int bhh_write_png(const char* path, int w, int h, int channels, const unsigned char* data, int level);

/* Per-frame exposure from the frames' lum_percentile (main.cpp:304-319): out[i] = key /
   exp(sum_j w_ij ln lp[j] / sum_j w_ij), w_ij = exp(-(i-j)^2 / 2 sigma^2) for |i-j| <= ceil(3 sigma)
   and lp[j] > 0 (sigma = 0: the frame alone); out[i] = 1 where no frame contributes. */
void bhh_smooth_exposure(const double* lp, int n, double key, double sigma, double* out);

/* HUD layout of frame `frame` of a shot (make_hud, main.cpp:195-300) in a W x H image for a font of
   cell height cell_h and advance `advance` px: up to glyph_cap glyphs and rect_cap rectangles, the
   numbers in *n_glyphs / *n_rects (may exceed the caps). Returns 0, or -1 with the error in err. */
int bhh_make_hud(const char* ini, const char* sets, int frame, int W, int H, int cell_h, float advance,
                 bh_hud_glyph* glyphs, int glyph_cap, int* n_glyphs, bh_hud_rect* rects, int rect_cap, int* n_rects,
                 char* err, size_t errlen);

/* Writes a font file for the HUD ("BHFONT1\n", int32 cell_w, cell_h, glyphs, baseline, float32
   advance, then the glyphs x cell_h x cell_w coverage bytes). Returns 0 or -1. */
int bhh_save_font(const char* path, const bh_font* font);

/* sizeof of the public structs of bhrt.h in this order: bh_scene, bh_tables, bh_frame, bh_settings,
   bh_trace_opts, bh_stats, bh_ray_out, bh_batch_frame, bh_batch_cfg, bh_mlp_train_cfg,
   bh_mlp_report, bh_mlp_bake_report, bh_post_params, bh_tone_params, bh_hud_glyph, bh_hud_rect,
   bh_font - so that a binding can check its struct layouts against the library's. No counterpart
   in the reference (its host code shares the headers). Writes min(n, count) sizes to out
   (nullable) and returns the count. */
int bhh_abi_sizes(size_t* out, int n);

/* offsetof of every field of the same structs, struct by struct in that order and field by field in
   declaration order (an array or nested struct counts as one field), so a binding can check each
   field's place, not only each struct's size. No counterpart in the reference. Writes min(n, count)
   offsets to out (nullable) and returns the count. */
int bhh_abi_offsets(size_t* out, int n);

/* Self-test rows (8 x 6: a, incl_deg, traced -, traced +, analytic -, analytic +) on `device`. */
int bhh_selftest(int device, double* rows, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif
#endif
