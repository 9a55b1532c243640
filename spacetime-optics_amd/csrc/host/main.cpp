/* bhrt - command line front end on the MI355X tracer (blackholes/src/main.cpp:374-626 of the
   reference, single GPU): bhrt shot.ini [--out DIR] [--set section.key=value ...] [--frames A:B]
   [--list] [--selftest]. Frames are rendered as the reference's Renderer::render does (adaptive
   sampling between render.spp and render.spp_max, then in radio mode the beam blur of post.blur x
   post.beam_uas and the frame's display statistics); the exposure of every frame follows from the
   smoothed statistics of all frames (post.exposure*, main.cpp:569-576), and each frame is tone
   mapped and written as out/frames/frame_NNNNN.png (main.cpp:578-603), with the HUD text overlay
   (make_hud, main.cpp:195-300) when hud.enabled (default 1) and a font file is given (hud.font, or
   --font FILE; the reference compiles its font in). With output.hdr = 1 the linear HDR frames (I, I_n1, I_n2, I_n>=3) are written as well:
   "BHF1" + int32 W, H, 4 + W*H*4 float32, top row first. A shot with a [batch] section (frames,
   spins, sizes, frame_dt_m) is traced in one persistent launch instead: frame i has spin
   spins[i % #spins], size sizes[(i / #spins) % #sizes] squared and time i * frame_dt_m. */
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <sstream>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/bhrt_host.h"
#include "config.h"
#include "hud.h"
#include "image_io.h"
#include "renderer.h"
#include "scene.h"
#include "shot.h"

namespace {

double wall() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

void usage() {
    std::fprintf(stderr,
                 "usage: bhrt shot.ini [--out DIR] [--set section.key=value ...] [--frames A:B] [--font FILE] [--list] "
                 "[--selftest]\n");
}

void mkdirs(const std::string& p) {
    std::string cur;
    for (size_t i = 0; i < p.size(); ++i) {
        cur += p[i];
        if (p[i] == '/' || i + 1 == p.size()) mkdir(cur.c_str(), 0755);
    }
}

bool write_frame(const std::string& path, const bh::HdrFrame& f) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    int32_t hdr[3] = {f.W, f.H, 4};
    bool ok = std::fwrite("BHF1", 1, 4, fp) == 4 && std::fwrite(hdr, 4, 3, fp) == 3 &&
              std::fwrite(f.px.data(), 4, f.px.size(), fp) == f.px.size();
    return std::fclose(fp) == 0 && ok;
}

// Phase 2 of the reference's frame loop (main.cpp:569-603): exposure from the smoothed display
// statistics of all frames, radio tone map, PNG per frame (and the HDR frame with output.hdr = 1).
// Visible-mode frames (vacuum geodesics and the checkerboard sky; the disk and the reference's ACES
// display map are not part of this tracer) are written as a linear grey PNG of the middle channel
// times post.exposure (default 1), and as HDR with output.hdr = 1.
// What the HUD layout needs besides the configuration and the shot
struct HudInputs {
    const bh::Font* font = nullptr;   // null: no overlay
    bh::SceneInfo si;
    bh_scene S{};
};

bool tone_and_write(const bh::Config& cfg, const bh::Shot& shot, bh::Renderer& R, bool radio,
                    const std::vector<int>& index, const std::vector<bh::HdrFrame>& frames, const std::string& out,
                    const HudInputs& hud) {
    const size_t n = frames.size();
    if (!radio) {
        const double gain = cfg.str("post.exposure", "auto") == "auto" ? 1.0 : cfg.num("post.exposure", 1.0);
        const bool hdr = cfg.flag("output.hdr", false);
        std::vector<uint8_t> grey;
        for (size_t i = 0; i < n; ++i) {
            const bh::HdrFrame& fr = frames[i];
            grey.resize((size_t)fr.W * fr.H);
            for (size_t p = 0; p < grey.size(); ++p) {
                const double v = std::min(1.0, std::max(0.0, gain * fr.px[4 * p + 1]));
                grey[p] = (uint8_t)std::lround(255.0 * v);
            }
            char name[64];
            std::snprintf(name, sizeof name, "/frames/frame_%05d.png", index[i]);
            if (!bh::write_png(out + name, fr.W, fr.H, 1, grey.data())) {
                std::fprintf(stderr, "bhrt: cannot write %s%s\n", out.c_str(), name);
                return false;
            }
            std::snprintf(name, sizeof name, "/frames/frame_%05d.bhf", index[i]);
            if (hdr && !write_frame(out + name, frames[i])) {
                std::fprintf(stderr, "bhrt: cannot write %s%s\n", out.c_str(), name);
                return false;
            }
        }
        return true;
    }
    std::vector<double> lp(n);
    for (size_t i = 0; i < n; ++i) lp[i] = frames[i].lum_percentile;
    const double key = cfg.num("post.exposure_key", 0.92);
    const double sigma = std::min(cfg.num("post.exposure_smooth", 4.0), (double)shot.frames.size() / 30.0);
    std::vector<double> ex = bh::smooth_exposure(lp, key, sigma);
    if (cfg.str("post.exposure", "auto") != "auto") ex.assign(n, key / cfg.num("post.exposure", 1.0));
    const bool hdr = cfg.flag("output.hdr", false);
    std::vector<uint8_t> rgb;
    for (size_t i = 0; i < n; ++i) {
        const int idx = index[i];
        const bh::ShotFrame& f = shot.frames[(size_t)std::min<size_t>((size_t)idx, shot.frames.size() - 1)];
        bh_tone_params tp{};
        tp.exposure = (float)(ex[i] * std::pow(2.0, f.exposure_ev));
        tp.colormap = cfg.str("post.colormap", "afmhot") == "inferno" ? 1 : 0;
        tp.gamma = (float)cfg.num("post.gamma", 1.0);
        tp.overlay = (float)f.overlay;
        tp.hide_direct = (float)f.hide_direct;
        tp.fade = (float)f.fade;
        tp.vignette = (float)cfg.num("post.vignette", 0.15);
        tp.grain = (float)cfg.num("post.grain", 0.006);
        tp.seed = 0x51ED27u + (uint32_t)idx * 7919u;
        bool ok;
        if (hud.font) {
            const bh_frame F = bh::make_frame(shot, f, hud.S.a, frames[i].W, frames[i].H, 1, 0);
            const bh::HudLayout L = bh::make_hud(cfg, shot, f, F, hud.si, hud.S, frames[i].W, frames[i].H,
                                                 hud.font->cell_h, hud.font->advance);
            const bh_font fv = hud.font->view();
            ok = R.tonemap(frames[i], tp, L.glyphs, L.rects, &fv, &rgb);
        } else {
            ok = R.tonemap(frames[i], tp, &rgb);
        }
        if (!ok) {
            std::fprintf(stderr, "bhrt: tone map of frame %d: %s\n", idx, R.error().c_str());
            return false;
        }
        char name[64];
        std::snprintf(name, sizeof name, "/frames/frame_%05d.png", idx);
        if (!bh::write_png(out + name, frames[i].W, frames[i].H, 3, rgb.data())) {
            std::fprintf(stderr, "bhrt: cannot write %s%s\n", out.c_str(), name);
            return false;
        }
        if (hdr) {
            std::snprintf(name, sizeof name, "/frames/frame_%05d.bhf", idx);
            if (!write_frame(out + name, frames[i])) {
                std::fprintf(stderr, "bhrt: cannot write %s%s\n", out.c_str(), name);
                return false;
            }
        }
    }
    return true;
}

int selftest() {
    std::printf("self-test: shadow edge from ray tracing vs analytic critical curve\n");
    std::vector<double> rows(48);
    char err[512] = {0};
    if (bhh_selftest(0, rows.data(), err, sizeof err) != BH_OK) {
        std::fprintf(stderr, "bhrt: self-test failed: %s\n", err);
        return 1;
    }
    for (int i = 0; i < 8; ++i) {
        const double* r = &rows[6 * i];
        std::printf("  a=%.2f incl=%4.0f  ray-traced edges %+.6f %+.6f   analytic %+.6f %+.6f  (M)\n", r[0], r[1], r[2],
                    r[3], r[4], r[5]);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    double t_start = wall();
    if (argc < 2) {
        usage();
        return 1;
    }
    if (std::string(argv[1]) == "--selftest") return selftest();
    bh::Config cfg;
    std::string err;
    if (!cfg.load(argv[1], &err)) {
        std::fprintf(stderr, "bhrt: %s\n", err.c_str());
        return 1;
    }
    bool list = false;
    int fa = 0, fb = -1;
    std::string font_path;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--set" && i + 1 < argc) cfg.apply_set(argv[++i]);
        else if (a == "--font" && i + 1 < argc) font_path = argv[++i];
        else if (a == "--out" && i + 1 < argc) cfg.set("output.dir", argv[++i]);
        else if (a == "--frames" && i + 1 < argc) std::sscanf(argv[++i], "%d:%d", &fa, &fb);
        else if (a == "--list") list = true;
        else {
            usage();
            return 1;
        }
    }
    bh_scene S;
    bh::SceneTables T;
    bh::SceneInfo si;
    bh::TracerOptions opt;
    bh::Shot shot;
    if (!bh::build_scene(cfg, &S, &T, &si, &opt, &err, !list) || !bh::build_shot(cfg, &shot, &err)) {
        std::fprintf(stderr, "bhrt: %s\n", err.c_str());
        return 1;
    }
    bh_settings rs{};
    rs.W = cfg.integer("camera.width", 1920);
    rs.H = cfg.integer("camera.height", 1080);
    rs.spp = cfg.integer("render.spp", 8);
    rs.spp_max = cfg.integer("render.spp_max", 128);
    rs.target_err = (float)cfg.num("render.target_error", 0.03);
    int nframes = (int)shot.frames.size();
    if (fb < 0 || fb > nframes) fb = nframes;
    fa = std::max(0, std::min(fa, fb));
    std::printf("bhrt | 1 GPU | shot '%s' (%s) %d frames %dx%d spp %d | %s%s\n", shot.name.c_str(), shot.type.c_str(),
                nframes, rs.W, rs.H, rs.spp, opt.precision == BH_PREC_MIXED ? "mixed" : "fp64",
                opt.emissivity == BH_EMIS_MLP ? " + mlp" : "");
    std::printf("  a = %.4f, r_h = %.4f M, r_isco = %.4f M, %s, M = %.3g Msun (GM/c^2 = %.4g km, %.3g uas)\n", S.a, S.rh,
                S.r_isco, S.mode == BH_MODE_RADIO ? "radio" : "visible", si.mass_msun, si.rg_cm * 1e-5, si.uas_per_m);
    if (list) {
        for (const bh::ShotFrame& f : shot.frames)
            std::printf("  frame %4d t=%6.2f r=%.4g incl=%.1f phi=%.1f fov=%.4g zoom=%.4g target=%.2f time=%.1f\n",
                        f.index, f.t, f.r, f.incl * 180 / bh::PI, f.phi * 180 / bh::PI,
                        f.fov_in_m ? f.fov : f.fov * 180 / bh::PI, f.zoom, f.target, f.time_m);
        return 0;
    }
    bh::Renderer R;
    if (bh_device_count() <= 0) {
        std::fprintf(stderr, "bhrt: no HIP device\n");
        return 1;
    }
    double t0 = wall();
    if (!R.init(0, S, T, rs, &err)) {
        std::fprintf(stderr, "bhrt: %s\n", err.c_str());
        return 1;
    }
    std::printf("  GPU init %.2f s\n", wall() - t0);
    if (shot.type == "render" && cfg.has("camera.target_psi_deg")) {
        double psi = cfg.num("camera.target_psi_deg", 0) * bh::PI / 180;
        bh_frame F0 = bh::make_frame(shot, shot.frames[0], S.a, rs.W, rs.H, rs.spp, 0);
        bh::FateFn fates = [&](const std::vector<double>& ph, std::vector<int>* out) {
            return R.vacuum_fates(ph, 1e-12, 200.0, 2000000, out);
        };
        std::vector<double> rho;
        if (!bh::find_critical_points(S.a, {F0}, {psi}, fates, &rho)) {
            std::fprintf(stderr, "bhrt: %s\n", R.error().c_str());
            return 1;
        }
        double r = rho[0] + cfg.num("camera.target_offset_m", 0.0);
        shot.target_x = r * std::cos(psi);
        shot.target_y = r * std::sin(psi);
        std::printf("  zoom target: critical curve at psi = %.1f deg, rho = %.12f M\n", psi * 180 / bh::PI, r);
    }
    if (S.mode == BH_MODE_RADIO && si.radio_flux_target_jy > 0) {
        double got = 0;
        if (!bh::normalize_flux(R, shot, shot.frames[0], si, &S, &got, &err)) {
            std::fprintf(stderr, "bhrt: %s\n", err.c_str());
            return 1;
        }
        std::printf("  flow normalised: n0 = %.4g cm^-3 for %.3g Jy (got %.3g Jy)\n", S.n0, si.radio_flux_target_jy, got);
    }
    std::string out = cfg.str("output.dir", "out/" + shot.name);
    mkdirs(out + "/frames");
    bh_trace_opts o{};
    o.precision = opt.precision;
    o.emissivity = opt.emissivity;
    o.jitter = opt.jitter;
    o.spp = rs.spp;
    if (opt.emissivity == BH_EMIS_MLP) {
        bh_mlp_report mr{};
        bh_mlp_bake_report br{};
        if (bh_mlp_train(R.ctx(), nullptr, &mr) != BH_OK || bh_mlp_bake(R.ctx(), 0, 0, 0, 0.f, &br) != BH_OK) {
            std::fprintf(stderr, "bhrt: surrogate: %s\n", R.error().c_str());
            return 1;
        }
        std::printf("  emissivity surrogate: trained in %.1f s (loss %.3g -> %.3g), %lld of %lld hot cells judged exact-only\n",
                    mr.ms_train * 1e-3, mr.loss_first, mr.loss_last, br.failed_hot_cells, br.hot_cells);
    }
    if (cfg.has("batch.frames")) {
        // one persistent launch over the [batch] frames
        auto list = [&](const std::string& key, const std::string& dflt) {
            std::vector<double> v;
            std::stringstream ss(cfg.str(key, dflt));
            std::string item;
            while (std::getline(ss, item, ','))
                if (!item.empty()) v.push_back(std::atof(item.c_str()));
            return v;
        };
        const int nb = cfg.integer("batch.frames", 1);
        const std::vector<double> spins = list("batch.spins", std::to_string(S.a));
        const std::vector<double> sizes = list("batch.sizes", std::to_string(rs.W));
        const double dt = cfg.num("batch.frame_dt_m", 0.0);
        if (spins.empty() || sizes.empty() || nb <= 0) {
            std::fprintf(stderr, "bhrt: bad [batch] section\n");
            return 1;
        }
        std::vector<bh_batch_frame> B((size_t)nb);
        for (int i = 0; i < nb; ++i) {
            bh::Config ci = cfg;
            ci.set("bh.spin", std::to_string(spins[(size_t)i % spins.size()]));
            bh_scene Si;
            bh::SceneTables Ti;
            bh::SceneInfo sii;
            bh::TracerOptions oi;
            if (!bh::build_scene(ci, &Si, &Ti, &sii, &oi, &err, false)) {
                std::fprintf(stderr, "bhrt: %s\n", err.c_str());
                return 1;
            }
            Si.n0 = S.n0;   // the normalised density of the base scene
            const int n = (int)sizes[((size_t)i / spins.size()) % sizes.size()];
            bh::ShotFrame sf = shot.frames[0];
            sf.time_m = i * dt;
            B[(size_t)i].scene = Si;
            B[(size_t)i].frame = bh::make_frame(shot, sf, Si.a, n, n, 1, 0x9E3779B9u * (uint32_t)(i + 1));
        }
        std::vector<bh::HdrFrame> imgs;
        bh_stats st{};
        double t1 = wall();
        if (!R.render_batch(B, o, &imgs, &st)) {
            std::fprintf(stderr, "bhrt: batch: %s\n", R.error().c_str());
            return 1;
        }
        const double tb = wall() - t1;
        std::vector<int> index((size_t)nb);
        const double percentile = cfg.num("post.exposure_percentile", 0.995);
        for (int i = 0; i < nb; ++i) {
            index[(size_t)i] = i;
            bh::HdrFrame& h = imgs[(size_t)i];
            bh_post_params pp{0.0, 0.f};
            if (bh_post_radio(R.ctx(), h.px.data(), h.W, h.H, &pp, percentile, &h.lum_percentile, &h.lum_max) != BH_OK) {
                std::fprintf(stderr, "bhrt: batch post: %s\n", R.error().c_str());
                return 1;
            }
        }
        if (!tone_and_write(cfg, shot, R, S.mode == BH_MODE_RADIO, index, imgs, out, HudInputs{})) return 1;
        std::printf("done: batch of %d frames, %.3g rays in %.1f ms kernel / %.2f s wall (%.1f Mrays/s)\n", nb,
                    (double)st.samples, st.ms_kernel, tb, st.samples / std::max(st.ms_kernel, 1e-9) * 1e-3);
        return 0;
    }
    long long samples = 0, limited = 0;
    double t_render = wall();
    // phase 1: render with the post step (main.cpp:519-564)
    const double percentile = cfg.num("post.exposure_percentile", 0.995);
    const double beam_px_uas = cfg.num("post.beam_uas", 20.0);
    std::vector<bh::HdrFrame> done;
    std::vector<int> index;
    for (int idx = fa; idx < fb; ++idx) {
        const bh::ShotFrame& sf = shot.frames[idx];
        bh_frame F = bh::make_frame(shot, sf, S.a, rs.W, rs.H, rs.spp, 0x9E3779B9u * (uint32_t)(idx + 1));
        bh_post_params pp{0.0, (float)sf.hide_direct};
        if (S.mode == BH_MODE_RADIO && sf.blur > 0 && sf.fov_in_m) {
            const double pix_uas = sf.fov / sf.zoom * si.uas_per_m / rs.W;
            pp.blur_sigma_px = sf.blur * beam_px_uas / 2.3548 / pix_uas;
        }
        bh::HdrFrame hdr;
        bh_stats st{};
        double t1 = wall();
        if (!R.render(F, o, pp, percentile, &hdr, &st)) {
            std::fprintf(stderr, "bhrt: frame %d: %s\n", idx, R.error().c_str());
            return 1;
        }
        samples += st.samples;
        limited += st.step_limited;
        std::printf("[gpu 0] %s frame %4d  %6.1f Msamples  %7.1f ms  (step-limited %lld)\n", shot.name.c_str(), idx,
                    st.samples * 1e-6, (wall() - t1) * 1e3, st.step_limited);
        std::fflush(stdout);
        done.push_back(std::move(hdr));
        index.push_back(idx);
    }
    double tr = wall() - t_render;
    // phase 2: exposure, tone map, PNG
    HudInputs hud;
    bh::Font font;
    if (S.mode == BH_MODE_RADIO && cfg.flag("hud.enabled", true)) {
        std::string fp = font_path.empty() ? cfg.str("hud.font", "") : font_path;
        if (!fp.empty() && fp[0] != '/' && font_path.empty()) fp = cfg.dir() + "/" + fp;
        if (fp.empty()) {
            std::printf("  HUD: no font (hud.font or --font); frames are written without the overlay\n");
        } else if (!bh::load_font(fp, &font, &err)) {
            std::fprintf(stderr, "bhrt: %s\n", err.c_str());
            return 1;
        } else {
            hud.font = &font;
            hud.si = si;
            hud.S = S;
        }
    }
    if (!tone_and_write(cfg, shot, R, S.mode == BH_MODE_RADIO, index, done, out, hud)) return 1;
    FILE* m = std::fopen((out + "/meta.json").c_str(), "w");
    if (m) {
        std::fprintf(m, "{\"shot\": \"%s\", \"frames\": %d, \"fps\": %g, \"width\": %d, \"height\": %d}\n", shot.name.c_str(),
                     nframes, shot.fps, rs.W, rs.H);
        std::fclose(m);
    }
    std::printf("done: %d frames, %.3g samples (%lld step-limited) in %.1f s render (%.3g Msamples/s), %.1f s total\n",
                fb - fa, (double)samples, limited, tr, samples / std::max(tr, 1e-9) * 1e-6, wall() - t_start);
    return 0;
}
