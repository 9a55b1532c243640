/* ref_driver.cpp - exposes functions of the reference's own host sources as a C ABI, so the golden
   vector script (tests/golden/make_golden.py) and the CPU tests can call them through ctypes.
   TEST INFRASTRUCTURE ONLY.

   Built by oracle/Makefile (target _ref/libref.so) against the reference sources where they lie:
   blackholes/src/{kerr.hpp, shot.cpp, scene_build.cpp, colour.cpp} and common/config.cpp, compiled
   unmodified with the reference's own flags (g++ -O3 -std=c++17, blackholes/Makefile:10). Nothing
   here replaces any part of the reference; it only calls it. The CUDA-only files (trace.cuh,
   emission.cuh, device_util.cuh, render.cu) do not compile with g++ and are not used. */
#include <cstring>
#include <string>

#include "config.hpp"
#include "kerr.hpp"
#include "scene_build.hpp"
#include "shot.hpp"

namespace {
void kp_out(const KerrPoint& k, double* o) {
    o[0] = k.r; o[1] = k.s; o[2] = k.c; o[3] = k.Sigma; o[4] = k.Delta; o[5] = k.A; o[6] = k.omega;
    o[7] = k.e_nu; o[8] = k.e_psi; o[9] = k.sqrtSigma; o[10] = k.sqrtSigmaDelta;
}
KerrPoint kp_in(const double* o) {
    KerrPoint k;
    k.r = o[0]; k.s = o[1]; k.c = o[2]; k.Sigma = o[3]; k.Delta = o[4]; k.A = o[5]; k.omega = o[6];
    k.e_nu = o[7]; k.e_psi = o[8]; k.sqrtSigma = o[9]; k.sqrtSigmaDelta = o[10];
    return k;
}
// frame out: r, th, ph, t, kp[11], F[3], R[3], U[3], beta[3], cx, cy, pix  (31 doubles)
void frame_out(const FrameGPU& F, double* o) {
    o[0] = F.r; o[1] = F.th; o[2] = F.ph; o[3] = F.t;
    kp_out(F.kp, o + 4);
    const dvec3* v[4] = {&F.F, &F.R, &F.U, &F.beta};
    for (int i = 0; i < 4; ++i) { o[15 + 3 * i] = v[i]->x; o[16 + 3 * i] = v[i]->y; o[17 + 3 * i] = v[i]->z; }
    o[27] = F.cx; o[28] = F.cy; o[29] = F.pix; o[30] = 0;
}
bool load_cfg(const char* path, const char* sets, Config* c) {
    std::string err;
    if (!c->load(path, &err)) return false;
    // sets: "key=value;key=value"
    std::string s = sets ? sets : "";
    size_t p = 0;
    while (p < s.size()) {
        size_t q = s.find(';', p);
        if (q == std::string::npos) q = s.size();
        std::string kv = s.substr(p, q - p);
        size_t eq = kv.find('=');
        if (eq != std::string::npos) c->set(trim(kv.substr(0, eq)), trim(kv.substr(eq + 1)));
        p = q + 1;
    }
    return true;
}
}  // namespace

extern "C" {

void ref_kerr_point(double a, double r, double th, double* o) { kp_out(kerr_point(a, r, th), o); }
double ref_horizon_radius(double a) { return horizon_radius(a); }
double ref_isco_radius(double a) { return isco_radius(a); }
void ref_camera_photon(double a, const double* kp, double phi0, double t0, const double* d, const double* b,
                       double* out) {
    Photon p = camera_photon(a, kp_in(kp), phi0, t0, dvec3{d[0], d[1], d[2]}, dvec3{b[0], b[1], b[2]});
    out[0] = p.lam; out[1] = p.eta; out[2] = p.E;
    for (int i = 0; i < 6; ++i) out[3 + i] = p.y[i];
}
double ref_radial_potential(double a, double lam, double eta, double r) { return radial_potential(a, lam, eta, r); }
double ref_polar_potential(double a, double lam, double eta, double u) { return polar_potential(a, lam, eta, u); }
void ref_geodesic_rhs(double a, double lam, double eta, const double* y, double* d) { geodesic_rhs(a, lam, eta, y, d); }
double ref_dp54_step(double a, double lam, double eta, const double* y, const double* k1, double h, double tol,
                     double* yout, double* k7) {
    return dp54_step(a, lam, eta, y, k1, h, tol, yout, k7);
}
void ref_project_radial(double a, double lam, double eta, double* y, double* k) { project_radial(a, lam, eta, y, k); }
double ref_initial_step(const double* y) { return initial_step(y); }
int ref_vacuum_fate(double a, const double* photon, double tol, double r_escape, int max_steps) {
    Photon p;
    p.lam = photon[0]; p.eta = photon[1]; p.E = photon[2];
    for (int i = 0; i < 6; ++i) p.y[i] = photon[3 + i];
    return vacuum_fate(a, p, tol, r_escape, max_steps);
}
int ref_critical_curve_point(double a, double theta_o, double rp, double* al, double* be) {
    return critical_curve_point(a, theta_o, rp, al, be) ? 1 : 0;
}

// Scene scalars from a shot .ini (+ "k=v;k=v" overrides): out[0..] as listed below; k2 table (512).
// out: a, rh, r_isco, tol, max_steps, r_escape, r_stop, step_frac, mode, flow_on, n0, n_index,
//      theta_index, theta0, h_r, beta_p, r_flow_max, nu_obs, rg_cm, flow_kep, flow_turb_amp,
//      k2_n, k2_l0, k2_dl, seed, distance_cm, uas_per_m, flux_target_jy, mass_msun
int ref_build_scene(const char* path, const char* sets, double* out, float* k2) {
    Config c;
    if (!load_cfg(path, sets, &c)) return 0;
    SceneGPU S;
    SceneTables T;
    SceneInfo I;
    std::string err;
    if (!build_scene(c, &S, &T, &I, &err)) return 0;
    double v[] = {S.a, S.rh, S.r_isco, S.tol, (double)S.max_steps, S.r_escape, S.r_stop, S.step_frac,
                  (double)S.mode, (double)S.flow_on, S.n0, S.n_index, S.theta_index, S.theta0, S.h_r, S.beta_p,
                  S.r_flow_max, S.nu_obs, S.rg_cm, S.flow_kep, (double)S.flow_turb_amp, (double)S.k2_n,
                  (double)S.k2_l0, (double)S.k2_dl, (double)S.seed, I.distance_cm, I.uas_per_m,
                  I.radio_flux_target_jy, I.mass_msun, (double)S.flow_turb_scale, (double)S.flow_tcorr};
    std::memcpy(out, v, sizeof v);
    for (int i = 0; i < S.k2_n; ++i) k2[i] = T.k2_ln[i];
    return 1;
}

// sky_rot of build_scene (scene_build.cpp:181-204), row-major, and sky_scale
int ref_sky_rot(const char* path, const char* sets, double* out) {
    Config c;
    if (!load_cfg(path, sets, &c)) return 0;
    SceneGPU S;
    SceneTables T;
    SceneInfo I;
    std::string err;
    if (!build_scene(c, &S, &T, &I, &err)) return 0;
    const dvec3 rows[3] = {S.sky_rot.r0, S.sky_rot.r1, S.sky_rot.r2};
    for (int i = 0; i < 3; ++i) {
        out[3 * i] = rows[i].x;
        out[3 * i + 1] = rows[i].y;
        out[3 * i + 2] = rows[i].z;
    }
    out[9] = S.sky_scale;
    return 1;
}

// Frames of a shot: per frame 14 values (t, r, incl, phi, roll, yaw, pitch, fov, fov_in_m, zoom,
// target, off_x, off_y, time_m). Returns the frame count (at most maxn written).
int ref_build_shot(const char* path, const char* sets, double* out, int maxn) {
    Config c;
    if (!load_cfg(path, sets, &c)) return -1;
    Shot sh;
    std::string err;
    if (!build_shot(c, &sh, &err)) return -1;
    int n = (int)sh.frames.size();
    for (int i = 0; i < n && i < maxn; ++i) {
        const ShotFrame& f = sh.frames[i];
        double v[20] = {f.t, f.r, f.incl, f.phi, f.roll, f.yaw, f.pitch, f.fov, f.fov_in_m ? 1.0 : 0.0,
                        f.zoom, f.target, f.off_x, f.off_y, f.time_m, f.blur, f.overlay, f.hide_direct,
                        f.exposure_ev, f.fade, f.hud};
        std::memcpy(out + 20 * i, v, sizeof v);
    }
    return n;
}

// Config lookup: copies the raw value of key (after include/overrides) into buf; returns 1 if found.
int ref_config_str(const char* path, const char* sets, const char* key, char* buf, int len) {
    Config c;
    if (!load_cfg(path, sets, &c)) return -1;
    if (!c.has(key)) return 0;
    std::string v = c.str(key);
    std::strncpy(buf, v.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return 1;
}

// make_frame (shot.cpp:116-146) for explicit camera parameters (angles in radians).
void ref_make_frame(double a, double r, double incl, double phi, double roll, double yaw, double pitch, double fov,
                    int fov_in_m, double zoom, double target, double tx, double ty, double off_x, double off_y,
                    double time_m, int orbit, int W, double* out) {
    Shot sh;
    sh.motion = orbit ? "orbit" : "static";
    sh.target_x = tx;
    sh.target_y = ty;
    ShotFrame f;
    f.r = r; f.incl = incl; f.phi = phi; f.roll = roll; f.yaw = yaw; f.pitch = pitch; f.fov = fov;
    f.fov_in_m = fov_in_m != 0; f.zoom = zoom; f.target = target; f.off_x = off_x; f.off_y = off_y;
    f.time_m = time_m;
    frame_out(make_frame(sh, f, a, W, 0), out);
}

// find_critical_point (shot.cpp:148-164) for a static camera with roll/yaw/pitch.
double ref_find_critical_point(double a, double r, double incl, double phi, double roll, double yaw, double pitch,
                               double psi) {
    Shot sh;
    sh.motion = "static";
    ShotFrame f;
    f.r = r; f.incl = incl; f.phi = phi; f.roll = roll; f.yaw = yaw; f.pitch = pitch;
    f.fov_in_m = true; f.fov = 30;
    return find_critical_point(sh, f, a, psi);
}

}  // extern "C"
