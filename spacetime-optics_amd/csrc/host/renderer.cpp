#include "renderer.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace bh {

Renderer::~Renderer() { bh_destroy(ctx_); }

std::string Renderer::error() const { return bh_last_error(ctx_); }

bool Renderer::init(int device, const bh_scene& scene, const SceneTables& tables, const bh_settings& rs,
                    std::string* err) {
    bh_destroy(ctx_);
    ctx_ = nullptr;
    S_ = scene;
    rs_ = rs;
    bh_tables t = tables.view();
    char buf[512] = {0};
    int rc = bh_create(device, &scene, &t, &rs, &ctx_, buf, sizeof buf);
    if (rc != BH_OK) {
        if (err) *err = buf;
        return false;
    }
    return true;
}

bool Renderer::update_scene(const bh_scene& s) {
    S_ = s;
    return ctx_ && bh_update_scene(ctx_, &s) == BH_OK;
}

bool Renderer::render(const bh_frame& f, const bh_trace_opts& opts, HdrFrame* out, bh_stats* st) {
    if (!ctx_) return false;
    out->W = f.W;
    out->H = f.H;
    out->px.resize((size_t)f.W * f.H * 4);
    if (rs_.spp_max > rs_.spp) return bh_render_adaptive(ctx_, &f, &opts, out->px.data(), nullptr, st) == BH_OK;
    return bh_render(ctx_, &f, &opts, out->px.data(), st) == BH_OK;
}

bool Renderer::render(const bh_frame& f, const bh_trace_opts& opts, const bh_post_params& post, double percentile,
                      HdrFrame* out, bh_stats* st) {
    if (!render(f, opts, out, st)) return false;
    bh_post_params pp = post;
    if (S_.mode != BH_MODE_RADIO) pp.blur_sigma_px = 0.0;   // the beam blur is a radio-mode filter
    return bh_post_radio(ctx_, out->px.data(), out->W, out->H, &pp, percentile, &out->lum_percentile,
                         &out->lum_max) == BH_OK;
}

bool Renderer::tonemap(const HdrFrame& in, const bh_tone_params& tp, std::vector<uint8_t>* rgb8) {
    if (!ctx_) return false;
    rgb8->resize((size_t)in.W * in.H * 3);
    return bh_tonemap(ctx_, in.px.data(), in.W, in.H, &tp, rgb8->data()) == BH_OK;
}

bool Renderer::tonemap(const HdrFrame& in, const bh_tone_params& tp, const std::vector<bh_hud_glyph>& glyphs,
                       const std::vector<bh_hud_rect>& rects, const bh_font* font, std::vector<uint8_t>* rgb8) {
    if (!ctx_) return false;
    rgb8->resize((size_t)in.W * in.H * 3);
    return bh_tonemap_hud(ctx_, in.px.data(), in.W, in.H, &tp, glyphs.data(), (int)glyphs.size(), rects.data(),
                          (int)rects.size(), font, rgb8->data()) == BH_OK;
}

std::vector<double> smooth_exposure(const std::vector<double>& lp, double key, double sigma) {
    const long n = (long)lp.size();
    std::vector<double> out((size_t)n, 1.0);
    const int r = (int)std::ceil(3 * sigma);
    for (long i = 0; i < n; ++i) {
        double ws = 0, acc = 0;
        for (int d = -r; d <= r; ++d) {
            const long j = i + d;
            if (j < 0 || j >= n || lp[(size_t)j] <= 0) continue;
            const double w = sigma > 0 ? std::exp(-0.5 * d * d / (sigma * sigma)) : (d == 0 ? 1.0 : 0.0);
            ws += w;
            acc += w * std::log(lp[(size_t)j]);
        }
        if (ws > 0) out[(size_t)i] = key / std::exp(acc / ws);
    }
    return out;
}

bool Renderer::render_batch(const std::vector<bh_batch_frame>& frames, const bh_trace_opts& opts,
                            std::vector<HdrFrame>* out, bh_stats* st) {
    if (!ctx_ || frames.empty()) return false;
    size_t total = 0;
    for (const bh_batch_frame& b : frames) total += (size_t)b.frame.W * b.frame.H;
    std::vector<float> px(total * 4);
    if (bh_trace_batch(ctx_, frames.data(), (int)frames.size(), &opts, nullptr, px.data(), st) != BH_OK) return false;
    out->resize(frames.size());
    size_t off = 0;
    for (size_t i = 0; i < frames.size(); ++i) {
        HdrFrame& h = (*out)[i];
        h.W = frames[i].frame.W;
        h.H = frames[i].frame.H;
        const size_t n = (size_t)h.W * h.H * 4;
        h.px.assign(px.begin() + off, px.begin() + off + n);
        off += n;
    }
    return true;
}

double Renderer::radio_flux(bh_frame f, int n, double pix_sr) {
    double jy = -1;
    if (!ctx_ || bh_radio_flux(ctx_, &f, n, pix_sr, &jy) != BH_OK) return -1;
    return jy;
}

bool Renderer::vacuum_fates(const std::vector<double>& photons, double tol, double r_escape, int max_steps,
                            std::vector<int>* fates) {
    int n = (int)(photons.size() / 9);
    fates->assign(n, 0);
    return ctx_ && bh_vacuum_fates(ctx_, photons.data(), n, tol, r_escape, max_steps, fates->data()) == BH_OK;
}

bool normalize_flux(Renderer& R, const Shot& shot, ShotFrame f0, const SceneInfo& info, bh_scene* S, double* got_jy,
                    std::string* err) {
    f0.zoom = 1;
    f0.target = 0;
    f0.off_x = f0.off_y = 0;
    f0.fov = info.normalize_fov_m;
    f0.fov_in_m = true;
    const int n = 160;
    bh_frame F = make_frame(shot, f0, S->a, n, n, 2, 1);
    double pix_sr = std::pow(f0.fov / n * info.rg_cm / info.distance_cm, 2.0);
    auto flux = [&]() {
        double jy = R.radio_flux(F, n, pix_sr);
        return jy;
    };
    double first = flux();
    if (first < 0) {
        if (err) *err = "radio_flux failed: " + R.error();
        return false;
    }
    double x0 = std::log(S->n0), f0v = std::log(std::max(first, 1e-30));
    const double target = std::log(info.radio_flux_target_jy);
    double x1 = x0 + (target - f0v) / 1.5, f1v = 0;
    for (int it = 0; it < 8; ++it) {
        S->n0 = std::exp(x1);
        R.update_scene(*S);
        double v = flux();
        if (v < 0) {
            if (err) *err = "radio_flux failed: " + R.error();
            return false;
        }
        f1v = std::log(std::max(v, 1e-30));
        if (std::fabs(f1v - target) < 0.01) break;
        double slope = (f1v - f0v) / (x1 - x0);
        if (!(slope > 0.2)) slope = 1.5;
        x0 = x1;
        f0v = f1v;
        x1 = x1 + (target - f1v) / slope;
    }
    S->n0 = std::exp(x1);
    R.update_scene(*S);
    if (got_jy) *got_jy = std::exp(f1v);
    return true;
}

bool selftest_rows(Renderer& R, std::vector<std::vector<double>>* rows, std::string* err) {
    const double spins[4] = {1e-6, 0.5, 0.94, 0.99}, incls[2] = {90.0, 17.0};
    rows->clear();
    for (double a : spins) {
        // the device context integrates with the scene's spin: one context per spin value
        bh_scene S{};
        S.a = a;
        S.rh = horizon_radius(a);
        S.r_isco = isco_radius(a);
        S.tol = 1e-12;
        S.max_steps = 2000000;
        if (!R.update_scene(S)) {
            if (err) *err = "update_scene failed";
            return false;
        }
        std::vector<bh_frame> frames;
        std::vector<double> psis;
        for (double incl_deg : incls) {
            Shot sh;
            sh.motion = "static";
            ShotFrame f;
            f.r = 1e7;
            f.incl = incl_deg * PI / 180;
            f.fov_in_m = true;
            f.fov = 30;
            bh_frame F = make_frame(sh, f, a, 1, 1, 1, 0);
            frames.push_back(F);
            psis.push_back(0.0);
            frames.push_back(F);
            psis.push_back(PI);
        }
        std::vector<double> rho;
        FateFn fates = [&](const std::vector<double>& ph, std::vector<int>* out) {
            return R.vacuum_fates(ph, 1e-12, 200.0, 2000000, out);
        };
        if (!find_critical_points(a, frames, psis, fates, &rho)) {
            if (err) *err = "vacuum fates failed: " + R.error();
            return false;
        }
        for (int k = 0; k < 2; ++k) {
            double incl = incls[k] * PI / 180;
            double amin = 1e9, amax = -1e9;
            double rp0 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-a))),
                   rp1 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(a)));
            for (int i = 0; i <= 200000; ++i) {
                double rp = rp0 + (rp1 - rp0) * i / 200000.0, al, be;
                if (critical_curve_point(a, incl, rp, &al, &be)) {
                    amin = std::min(amin, al);
                    amax = std::max(amax, al);
                }
            }
            rows->push_back({a, incls[k], -rho[2 * k + 1], rho[2 * k], amin, amax});
        }
    }
    return true;
}

}  // namespace bh
