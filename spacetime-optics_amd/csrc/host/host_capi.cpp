/* host_capi.cpp - C entry points of the host layer (include/bhrt_host.h). */
#include <cmath>
#include <cstring>
#include <string>

#include <cstddef>
#include "../../../include/bhrt_host.h"
#include "config.h"
#include "hud.h"
#include "image_io.h"
#include "renderer.h"
#include "scene.h"
#include "shot.h"

namespace {

void copy_err(const std::string& m, char* err, size_t len) {
    if (err && len) {
        std::strncpy(err, m.c_str(), len - 1);
        err[len - 1] = 0;
    }
}

bool load(const char* ini, const char* sets, bh::Config* c, std::string* err) {
    if (!ini || !c->load(ini, err)) return false;
    std::string s = sets ? sets : "";
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find(';', p);
        if (q == std::string::npos) q = s.size();
        c->apply_set(s.substr(p, q - p));
        p = q + 1;
    }
    return true;
}

}  // namespace

extern "C" {

int bhh_config_str(const char* ini, const char* sets, const char* key, char* buf, int len) {
    bh::Config c;
    std::string err;
    if (!load(ini, sets, &c, &err)) {
        copy_err(err, buf, len);
        return -1;
    }
    if (!c.has(key)) return 0;
    copy_err(c.str(key), buf, len);
    return 1;
}

int bhh_build_scene(const char* ini, const char* sets, bh_scene* scene, float* k2, float* k2_l0, float* k2_dl,
                    double* info, bh_trace_opts* opts, char* err, size_t errlen) {
    bh::Config c;
    std::string e;
    if (!load(ini, sets, &c, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_ARG;
    }
    bh::SceneTables T;
    bh::SceneInfo I;
    bh::TracerOptions O;
    bh_scene S;
    if (!bh::build_scene(c, &S, &T, &I, &O, &e, false)) {
        copy_err(e, err, errlen);
        return BH_ERR_ARG;
    }
    if (scene) *scene = S;
    if (k2) std::memcpy(k2, T.k2_ln.data(), sizeof(float) * T.k2_ln.size());
    if (k2_l0) *k2_l0 = T.k2_l0;
    if (k2_dl) *k2_dl = T.k2_dl;
    if (info) {
        double v[6] = {I.mass_msun, I.distance_cm, I.rg_cm, I.uas_per_m, I.radio_flux_target_jy, I.normalize_fov_m};
        std::memcpy(info, v, sizeof v);
    }
    if (opts) {
        std::memset(opts, 0, sizeof(*opts));
        opts->precision = O.precision;
        opts->emissivity = O.emissivity;
        opts->jitter = O.jitter;
        opts->spp = c.integer("render.spp", 8);
    }
    return BH_OK;
}

void bhh_turbulence(int n, int nt, double half, double sigma, double corr_m, double corr_t, double dt, uint32_t seed,
                    float* out) {
    std::vector<float> v;
    bh::make_turbulence(n, nt, half, sigma, corr_m, corr_t, dt, seed, &v);
    std::memcpy(out, v.data(), sizeof(float) * v.size());
}

int bhh_build_shot(const char* ini, const char* sets, double* out, int maxn) {
    bh::Config c;
    std::string e;
    if (!load(ini, sets, &c, &e)) return -1;
    bh::Shot sh;
    if (!bh::build_shot(c, &sh, &e)) return -1;
    int n = (int)sh.frames.size();
    for (int i = 0; i < n && i < maxn; ++i) {
        const bh::ShotFrame& f = sh.frames[i];
        double v[BHH_SHOT_VALUES] = {f.t, f.r, f.incl, f.phi, f.roll, f.yaw, f.pitch, f.fov, f.fov_in_m ? 1.0 : 0.0,
                                     f.zoom, f.target, f.off_x, f.off_y, f.time_m, f.blur, f.overlay, f.hide_direct,
                                     f.exposure_ev, f.fade, f.hud};
        std::memcpy(out + BHH_SHOT_VALUES * (size_t)i, v, sizeof v);
    }
    return n;
}

void bhh_make_frame(double a, double r, double incl, double phi, double roll, double yaw, double pitch, double fov,
                    int fov_in_m, double zoom, double target, double target_x, double target_y, double off_x,
                    double off_y, double time_m, int orbit, int W, int H, int spp, uint32_t seed, bh_frame* out) {
    bh::Shot sh;
    sh.motion = orbit ? "orbit" : "static";
    sh.target_x = target_x;
    sh.target_y = target_y;
    bh::ShotFrame f;
    f.r = r; f.incl = incl; f.phi = phi; f.roll = roll; f.yaw = yaw; f.pitch = pitch; f.fov = fov;
    f.fov_in_m = fov_in_m != 0; f.zoom = zoom; f.target = target; f.off_x = off_x; f.off_y = off_y;
    f.time_m = time_m;
    *out = bh::make_frame(sh, f, a, W, H, spp, seed);
}

int bhh_find_critical_point(bh_ctx* ctx, const bh_frame* frame, double psi, double* rho) {
    if (!ctx || !frame || !rho) return BH_ERR_ARG;
    bh::FateFn fates = [&](const std::vector<double>& ph, std::vector<int>* out) {
        int n = (int)(ph.size() / 9);
        out->assign(n, 0);
        return bh_vacuum_fates(ctx, ph.data(), n, 1e-12, 200.0, 2000000, out->data()) == BH_OK;
    };
    std::vector<double> r;
    bh_scene S;
    bh_get_scene(ctx, &S);
    if (!bh::find_critical_points(S.a, {*frame}, {psi}, fates, &r)) return BH_ERR_HIP;
    *rho = r[0];
    return BH_OK;
}

int bhh_create_from_ini(const char* ini, const char* sets, int device, int normalize, int frame_index, bh_ctx** ctx,
                        bh_scene* scene, bh_frame* frame, bh_trace_opts* opts, double* jy, char* err, size_t errlen) {
    if (!ctx) return BH_ERR_ARG;
    *ctx = nullptr;
    bh::Config c;
    std::string e;
    if (!load(ini, sets, &c, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_ARG;
    }
    bh::SceneTables T;
    bh::SceneInfo I;
    bh::TracerOptions O;
    bh_scene S;
    bh::Shot shot;
    if (!bh::build_scene(c, &S, &T, &I, &O, &e) || !bh::build_shot(c, &shot, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_ARG;
    }
    if (frame_index < 0 || frame_index >= (int)shot.frames.size()) {
        copy_err("frame index out of range", err, errlen);
        return BH_ERR_ARG;
    }
    bh_settings rs{};
    rs.W = c.integer("camera.width", 1920);
    rs.H = c.integer("camera.height", 1080);
    rs.spp = c.integer("render.spp", 8);
    rs.spp_max = c.integer("render.spp_max", 128);
    rs.target_err = (float)c.num("render.target_error", 0.03);
    bh::Renderer R;
    if (!R.init(device, S, T, rs, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_NODEVICE;
    }
    // zoom target on the critical curve (main.cpp:421-428)
    if (shot.type == "render" && c.has("camera.target_psi_deg")) {
        double psi = c.num("camera.target_psi_deg", 0) * bh::PI / 180;
        bh_frame F0 = bh::make_frame(shot, shot.frames[0], S.a, rs.W, rs.H, rs.spp, 0);
        bh_scene Sv = S;   // the search runs with its own tolerance; the scene's spin is what matters
        bh::FateFn fates = [&](const std::vector<double>& ph, std::vector<int>* out) {
            return R.vacuum_fates(ph, 1e-12, 200.0, 2000000, out);
        };
        std::vector<double> rho;
        if (!bh::find_critical_points(Sv.a, {F0}, {psi}, fates, &rho)) {
            copy_err("critical point search failed: " + R.error(), err, errlen);
            return BH_ERR_HIP;
        }
        double r = rho[0] + c.num("camera.target_offset_m", 0.0);
        shot.target_x = r * std::cos(psi);
        shot.target_y = r * std::sin(psi);
    }
    double got = 0;
    if (normalize && S.mode == BH_MODE_RADIO && I.radio_flux_target_jy > 0) {
        if (!bh::normalize_flux(R, shot, shot.frames[0], I, &S, &got, &e)) {
            copy_err(e, err, errlen);
            return BH_ERR_HIP;
        }
    }
    if (jy) *jy = got;
    if (scene) *scene = S;
    if (frame)
        *frame = bh::make_frame(shot, shot.frames[frame_index], S.a, rs.W, rs.H, rs.spp,
                                0x9E3779B9u * (uint32_t)(frame_index + 1));
    if (opts) {
        std::memset(opts, 0, sizeof(*opts));
        opts->precision = O.precision;
        opts->emissivity = O.emissivity;
        opts->jitter = O.jitter;
        opts->spp = rs.spp;
    }
    *ctx = R.release();
    return BH_OK;
}

int bhh_write_png(const char* path, int w, int h, int channels, const unsigned char* data, int level) {
    if (!path) return -1;
    return bh::write_png(path, w, h, channels, data, level) ? 0 : -1;
}

void bhh_smooth_exposure(const double* lp, int n, double key, double sigma, double* out) {
    if (!lp || !out || n <= 0) return;
    const std::vector<double> e = bh::smooth_exposure(std::vector<double>(lp, lp + n), key, sigma);
    for (int i = 0; i < n; ++i) out[i] = e[(size_t)i];
}

int bhh_make_hud(const char* ini, const char* sets, int frame, int W, int H, int cell_h, float advance,
                 bh_hud_glyph* glyphs, int glyph_cap, int* n_glyphs, bh_hud_rect* rects, int rect_cap, int* n_rects,
                 char* err, size_t errlen) {
    bh::Config c;
    std::string e;
    if (W <= 0 || H <= 0 || cell_h <= 0 || !(advance > 0.f)) {
        copy_err("bad image size or font metrics", err, errlen);
        return -1;
    }
    bh_scene S;
    bh::SceneTables T;
    bh::SceneInfo si;
    bh::TracerOptions opt;
    bh::Shot sh;
    if (!load(ini, sets, &c, &e) || !bh::build_scene(c, &S, &T, &si, &opt, &e, false) || !bh::build_shot(c, &sh, &e)) {
        copy_err(e, err, errlen);
        return -1;
    }
    if (frame < 0 || frame >= (int)sh.frames.size()) {
        copy_err("frame index outside the shot", err, errlen);
        return -1;
    }
    const bh::ShotFrame& f = sh.frames[(size_t)frame];
    const bh_frame F = bh::make_frame(sh, f, S.a, W, H, 1, 0);
    const bh::HudLayout L = bh::make_hud(c, sh, f, F, si, S, W, H, cell_h, advance);
    if (n_glyphs) *n_glyphs = (int)L.glyphs.size();
    if (n_rects) *n_rects = (int)L.rects.size();
    for (int i = 0; i < (int)L.glyphs.size() && i < glyph_cap && glyphs; ++i) glyphs[i] = L.glyphs[(size_t)i];
    for (int i = 0; i < (int)L.rects.size() && i < rect_cap && rects; ++i) rects[i] = L.rects[(size_t)i];
    return 0;
}

int bhh_abi_sizes(size_t* out, int n) {
    const size_t sz[] = {sizeof(bh_scene),        sizeof(bh_tables),      sizeof(bh_frame),
                         sizeof(bh_settings),     sizeof(bh_trace_opts),  sizeof(bh_stats),
                         sizeof(bh_ray_out),      sizeof(bh_batch_frame), sizeof(bh_batch_cfg),
                         sizeof(bh_mlp_train_cfg), sizeof(bh_mlp_report), sizeof(bh_mlp_bake_report),
                         sizeof(bh_post_params),  sizeof(bh_tone_params), sizeof(bh_hud_glyph),
                         sizeof(bh_hud_rect),     sizeof(bh_font)};
    const int count = (int)(sizeof sz / sizeof sz[0]);
    for (int i = 0; out && i < n && i < count; ++i) out[i] = sz[i];
    return count;
}

int bhh_abi_offsets(size_t* out, int n) {
#define BH_OFF(T, f) offsetof(T, f)
    const size_t off[] = {
        // bh_scene
        BH_OFF(bh_scene, a), BH_OFF(bh_scene, rh), BH_OFF(bh_scene, r_isco), BH_OFF(bh_scene, tol),
        BH_OFF(bh_scene, max_steps), BH_OFF(bh_scene, r_escape), BH_OFF(bh_scene, r_stop), BH_OFF(bh_scene, step_frac),
        BH_OFF(bh_scene, mode), BH_OFF(bh_scene, flow_on), BH_OFF(bh_scene, n0), BH_OFF(bh_scene, n_index),
        BH_OFF(bh_scene, theta_index), BH_OFF(bh_scene, theta0), BH_OFF(bh_scene, h_r), BH_OFF(bh_scene, beta_p),
        BH_OFF(bh_scene, r_flow_max), BH_OFF(bh_scene, nu_obs), BH_OFF(bh_scene, rg_cm), BH_OFF(bh_scene, flow_kep),
        BH_OFF(bh_scene, flow_turb_amp), BH_OFF(bh_scene, seed), BH_OFF(bh_scene, flow_turb_scale),
        BH_OFF(bh_scene, flow_tcorr), BH_OFF(bh_scene, sky_on), BH_OFF(bh_scene, sky_nu), BH_OFF(bh_scene, sky_nv),
        BH_OFF(bh_scene, sky_scale), BH_OFF(bh_scene, sky_rot),
        // bh_tables
        BH_OFF(bh_tables, k2_ln), BH_OFF(bh_tables, k2_n), BH_OFF(bh_tables, k2_l0), BH_OFF(bh_tables, k2_dl),
        BH_OFF(bh_tables, turb), BH_OFF(bh_tables, turb_n), BH_OFF(bh_tables, turb_nt), BH_OFF(bh_tables, turb_half),
        BH_OFF(bh_tables, turb_t0), BH_OFF(bh_tables, turb_dt), BH_OFF(bh_tables, turb_model), BH_OFF(bh_tables, ts_dims),
        // bh_frame
        BH_OFF(bh_frame, r), BH_OFF(bh_frame, th), BH_OFF(bh_frame, ph), BH_OFF(bh_frame, t), BH_OFF(bh_frame, kp),
        BH_OFF(bh_frame, F), BH_OFF(bh_frame, R), BH_OFF(bh_frame, U), BH_OFF(bh_frame, beta), BH_OFF(bh_frame, cx),
        BH_OFF(bh_frame, cy), BH_OFF(bh_frame, pix), BH_OFF(bh_frame, W), BH_OFF(bh_frame, H), BH_OFF(bh_frame, spp),
        BH_OFF(bh_frame, seed),
        // bh_settings
        BH_OFF(bh_settings, W), BH_OFF(bh_settings, H), BH_OFF(bh_settings, spp), BH_OFF(bh_settings, spp_max),
        BH_OFF(bh_settings, target_err),
        // bh_trace_opts
        BH_OFF(bh_trace_opts, precision), BH_OFF(bh_trace_opts, emissivity), BH_OFF(bh_trace_opts, jitter),
        BH_OFF(bh_trace_opts, spp), BH_OFF(bh_trace_opts, tol_far), BH_OFF(bh_trace_opts, r_near),
        BH_OFF(bh_trace_opts, r_emit), BH_OFF(bh_trace_opts, far_fp32), BH_OFF(bh_trace_opts, full_flow),
        BH_OFF(bh_trace_opts, engine), BH_OFF(bh_trace_opts, tol_t), BH_OFF(bh_trace_opts, relax_after),
        BH_OFF(bh_trace_opts, tol_relax), BH_OFF(bh_trace_opts, relax_cap),
        // bh_stats
        BH_OFF(bh_stats, ms_trace), BH_OFF(bh_stats, ms_kernel), BH_OFF(bh_stats, samples), BH_OFF(bh_stats, step_limited),
        BH_OFF(bh_stats, step_attempts), BH_OFF(bh_stats, steps_accepted), BH_OFF(bh_stats, flow_steps),
        BH_OFF(bh_stats, fp64_steps), BH_OFF(bh_stats, flux_sum),
        // bh_ray_out
        BH_OFF(bh_ray_out, y), BH_OFF(bh_ray_out, th_inf), BH_OFF(bh_ray_out, ph_inf), BH_OFF(bh_ray_out, I),
        BH_OFF(bh_ray_out, trans), BH_OFF(bh_ray_out, psi), BH_OFF(bh_ray_out, fate), BH_OFF(bh_ray_out, ncross),
        BH_OFF(bh_ray_out, order), BH_OFF(bh_ray_out, steps), BH_OFF(bh_ray_out, accepted),
        BH_OFF(bh_ray_out, flow_steps), BH_OFF(bh_ray_out, sky_u), BH_OFF(bh_ray_out, sky_v),
        // bh_batch_frame
        BH_OFF(bh_batch_frame, scene), BH_OFF(bh_batch_frame, frame),
        // bh_batch_cfg
        BH_OFF(bh_batch_cfg, block), BH_OFF(bh_batch_cfg, grid_per_cu), BH_OFF(bh_batch_cfg, interval),
        BH_OFF(bh_batch_cfg, refill_min), BH_OFF(bh_batch_cfg, use_graph), BH_OFF(bh_batch_cfg, waves),
        BH_OFF(bh_batch_cfg, order), BH_OFF(bh_batch_cfg, prio_age),
        // bh_mlp_train_cfg
        BH_OFF(bh_mlp_train_cfg, n_samples), BH_OFF(bh_mlp_train_cfg, batch), BH_OFF(bh_mlp_train_cfg, epochs),
        BH_OFF(bh_mlp_train_cfg, lr), BH_OFF(bh_mlp_train_cfg, lr_min), BH_OFF(bh_mlp_train_cfg, hot_weight),
        BH_OFF(bh_mlp_train_cfg, seed),
        // bh_mlp_report
        BH_OFF(bh_mlp_report, ms_sample), BH_OFF(bh_mlp_report, ms_train), BH_OFF(bh_mlp_report, steps),
        BH_OFF(bh_mlp_report, loss_first), BH_OFF(bh_mlp_report, loss_last), BH_OFF(bh_mlp_report, train_flop),
        // bh_mlp_bake_report
        BH_OFF(bh_mlp_bake_report, ms_bake), BH_OFF(bh_mlp_bake_report, ms_judge), BH_OFF(bh_mlp_bake_report, cells),
        BH_OFF(bh_mlp_bake_report, hot_cells), BH_OFF(bh_mlp_bake_report, failed_hot_cells),
        // bh_post_params
        BH_OFF(bh_post_params, blur_sigma_px), BH_OFF(bh_post_params, hide_direct),
        // bh_tone_params
        BH_OFF(bh_tone_params, colormap), BH_OFF(bh_tone_params, exposure), BH_OFF(bh_tone_params, gamma),
        BH_OFF(bh_tone_params, overlay), BH_OFF(bh_tone_params, hide_direct), BH_OFF(bh_tone_params, fade),
        BH_OFF(bh_tone_params, vignette), BH_OFF(bh_tone_params, grain), BH_OFF(bh_tone_params, seed),
        // bh_hud_glyph
        BH_OFF(bh_hud_glyph, x), BH_OFF(bh_hud_glyph, y), BH_OFF(bh_hud_glyph, scale), BH_OFF(bh_hud_glyph, glyph),
        BH_OFF(bh_hud_glyph, rgba),
        // bh_hud_rect
        BH_OFF(bh_hud_rect, x0), BH_OFF(bh_hud_rect, y0), BH_OFF(bh_hud_rect, x1), BH_OFF(bh_hud_rect, y1),
        BH_OFF(bh_hud_rect, rgba), BH_OFF(bh_hud_rect, layer),
        // bh_font
        BH_OFF(bh_font, atlas), BH_OFF(bh_font, cell_w), BH_OFF(bh_font, cell_h), BH_OFF(bh_font, glyphs),
        BH_OFF(bh_font, advance), BH_OFF(bh_font, baseline)};
#undef BH_OFF
    const int count = (int)(sizeof off / sizeof off[0]);
    for (int i = 0; out && i < n && i < count; ++i) out[i] = off[i];
    return count;
}

int bhh_save_font(const char* path, const bh_font* font) {
    if (!path || !font || !font->atlas || font->cell_w < 1 || font->cell_h < 1 || font->glyphs < 1) return -1;
    bh::Font f;
    f.cell_w = font->cell_w;
    f.cell_h = font->cell_h;
    f.glyphs = font->glyphs;
    f.baseline = font->baseline;
    f.advance = font->advance;
    f.atlas.assign(font->atlas, font->atlas + (size_t)font->cell_w * font->cell_h * font->glyphs);
    std::string e;
    return bh::save_font(path, f, &e) ? 0 : -1;
}

int bhh_selftest(int device, double* rows, char* err, size_t errlen) {
    bh::Renderer R;
    bh_scene S{};
    S.a = 0.5;
    S.rh = bh::horizon_radius(S.a);
    S.r_isco = bh::isco_radius(S.a);
    S.tol = 1e-12;
    S.max_steps = 2000000;
    bh::SceneTables T;
    bh_settings rs{};
    std::string e;
    if (!R.init(device, S, T, rs, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_NODEVICE;
    }
    std::vector<std::vector<double>> out;
    if (!bh::selftest_rows(R, &out, &e)) {
        copy_err(e, err, errlen);
        return BH_ERR_HIP;
    }
    for (size_t i = 0; i < out.size(); ++i)
        for (int k = 0; k < 6; ++k) rows[6 * i + k] = out[i][k];
    return BH_OK;
}

}  // extern "C"
