#include "shot.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "scene.h"

namespace bh {

namespace {

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalize(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

// Rodrigues rotation (shot.cpp:43-46)
V3 rotate(V3 v, V3 axis, double ang) {
    double cs = std::cos(ang), sn = std::sin(ang);
    return v * cs + cross(axis, v) * sn + axis * (dot(axis, v) * (1 - cs));
}

bool parse_track(const std::string& s, bool log, Track* tr, std::string* err) {
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (item.empty()) continue;
        size_t c = item.find(':');
        if (c == std::string::npos) {
            if (err) *err = "track entry '" + item + "' needs time:value";
            return false;
        }
        double t = std::strtod(item.substr(0, c).c_str(), nullptr);
        double v = std::strtod(item.substr(c + 1).c_str(), nullptr);
        tr->t.push_back(t);
        tr->v.push_back(log ? std::log(v) : v);
    }
    tr->log = log;
    return !tr->t.empty();
}

// value of `key` at shot time `time`: the [track] entry if it parses, else the constant (shot.cpp:29-41)
double param(const Config& c, const std::string& key, double def, double time, bool log) {
    const std::string tk = "track." + key;
    if (c.has(tk)) {
        Track tr;
        std::string err;
        if (parse_track(c.str(tk), log, &tr, &err)) return tr.value(time);
        std::fprintf(stderr, "warning: %s\n", err.c_str());
    }
    return c.num(key, def);
}

}  // namespace

// Monotone cubic Hermite with Fritsch-Carlson tangents and zero end tangents (shot.cpp:49-71)
double Track::value(double time) const {
    const size_t n = t.size();
    double out;
    if (n == 1 || time <= t[0]) out = v[0];
    else if (time >= t[n - 1]) out = v[n - 1];
    else {
        size_t i = 0;
        while (i + 2 < n && time > t[i + 1]) ++i;
        auto secant = [&](size_t k) { return (v[k + 1] - v[k]) / (t[k + 1] - t[k]); };
        auto tangent = [&](size_t k) {
            if (k == 0 || k == n - 1) return 0.0;
            double s0 = secant(k - 1), s1 = secant(k);
            if (s0 * s1 <= 0) return 0.0;
            double h0 = t[k] - t[k - 1], h1 = t[k + 1] - t[k];
            return 3.0 * (h0 + h1) / ((2.0 * h1 + h0) / s0 + (h1 + 2.0 * h0) / s1);
        };
        double h = t[i + 1] - t[i], s = (time - t[i]) / h;
        double h00 = (1 + 2 * s) * (1 - s) * (1 - s), h10 = s * (1 - s) * (1 - s), h01 = s * s * (3 - 2 * s),
               h11 = s * s * (s - 1);
        out = h00 * v[i] + h10 * h * tangent(i) + h01 * v[i + 1] + h11 * h * tangent(i + 1);
    }
    return log ? std::exp(out) : out;
}

// shot.cpp:73-114 (the camera and time parameters; post/HUD tracks are out of scope)
bool build_shot(const Config& c, Shot* shot, std::string* err) {
    shot->name = c.str("shot.name", "shot");
    shot->type = c.str("shot.type", "render");
    shot->duration = c.num("shot.duration", 5);
    shot->fps = c.num("shot.fps", 30);
    shot->motion = c.str("camera.motion", "static");
    if (!(shot->fps > 0)) {
        if (err) *err = "shot.fps must be > 0";
        return false;
    }
    int n = std::max(1, (int)std::lround(shot->duration * shot->fps));
    if (c.has("shot.frames")) n = c.integer("shot.frames", n);
    const double D2R = PI / 180;
    shot->frames.clear();
    for (int i = 0; i < n; ++i) {
        ShotFrame f;
        f.index = i;
        f.t = i / shot->fps;
        const double t = f.t;
        f.r = param(c, "camera.r", 1e6, t, true);
        f.incl = param(c, "camera.incl", 80, t, false) * D2R;
        f.phi = param(c, "camera.phi", 0, t, false) * D2R;
        f.roll = param(c, "camera.roll", 0, t, false) * D2R;
        f.yaw = param(c, "camera.yaw", 0, t, false) * D2R;
        f.pitch = param(c, "camera.pitch", 0, t, false) * D2R;
        f.fov_in_m = c.has("camera.fov_m") || c.has("track.camera.fov_m");
        f.fov = f.fov_in_m ? param(c, "camera.fov_m", 30, t, true) : param(c, "camera.fov_deg", 40, t, true) * D2R;
        f.zoom = param(c, "camera.zoom", 1, t, true);
        f.target = param(c, "camera.target", 0, t, false);
        f.off_x = param(c, "camera.offset_x_m", 0, t, false);
        f.off_y = param(c, "camera.offset_y_m", 0, t, false);
        f.time_m = param(c, "time.t_m", 0, t, false);
        // post-processing timeline (shot.cpp:89-99): beam blur, order tint, direct-image removal,
        // exposure offset, fades, HUD opacity
        f.blur = param(c, "post.blur", 0, t, false);
        f.overlay = param(c, "post.overlay", 0, t, false);
        f.hide_direct = param(c, "post.hide_direct", 0, t, false);
        f.exposure_ev = param(c, "post.exposure_ev", 0, t, false);
        const double fade_in = c.num("shot.fade_in", 0.0), fade_out = c.num("shot.fade_out", 0.0);
        double fade = param(c, "post.fade", 1, t, false);
        if (fade_in > 0) fade *= std::min(1.0, t / fade_in);
        if (fade_out > 0) fade *= std::min(1.0, (shot->duration - t) / fade_out);
        f.fade = std::max(0.0, fade);
        f.hud = param(c, "hud.alpha", 1, t, false);
        shot->frames.push_back(f);
    }
    return true;
}

bh_frame make_frame(const Shot& shot, const ShotFrame& f, double a, int W, int H, int spp, uint32_t seed) {
    bh_frame F;
    std::memset(&F, 0, sizeof F);
    F.r = f.r;
    F.th = f.incl;
    F.ph = f.phi;
    F.t = f.time_m;
    // KerrPoint at the camera (kerr.hpp:27-42)
    {
        double r = f.r, s = std::sin(f.incl), c = std::cos(f.incl);
        double a2 = a * a, r2 = r * r;
        double Sigma = r2 + a2 * c * c;
        double Delta = r2 - 2.0 * r + a2;
        double A = (r2 + a2) * (r2 + a2) - a2 * Delta * s * s;
        double kp[11] = {r, s, c, Sigma, Delta, A, 2.0 * a * r / A, std::sqrt(Sigma * Delta / A),
                         std::sqrt(A / Sigma) * s, std::sqrt(Sigma), std::sqrt(Sigma * Delta)};
        std::memcpy(F.kp, kp, sizeof kp);
    }
    // Looking at the hole: forward = -e_r, up = -e_theta, right = e_phi
    V3 Fw = {-1, 0, 0}, Up = {0, -1, 0}, Rt = {0, 0, 1};
    if (f.yaw != 0) {
        Fw = rotate(Fw, Up, -f.yaw);
        Rt = rotate(Rt, Up, -f.yaw);
    }
    if (f.pitch != 0) {
        Fw = rotate(Fw, Rt, f.pitch);
        Up = rotate(Up, Rt, f.pitch);
    }
    if (f.roll != 0) {
        Rt = rotate(Rt, Fw, f.roll);
        Up = rotate(Up, Fw, f.roll);
    }
    const V3* src[3] = {&Fw, &Rt, &Up};
    double* dst[3] = {F.F, F.R, F.U};
    for (int i = 0; i < 3; ++i) {
        dst[i][0] = src[i]->x;
        dst[i][1] = src[i]->y;
        dst[i][2] = src[i]->z;
    }
    double width = f.fov_in_m ? f.fov / f.r : 2.0 * std::tan(0.5 * f.fov);
    F.pix = width / f.zoom / W;
    if (f.fov_in_m) {
        F.cx = (f.off_x + f.target * shot.target_x) / f.r;
        F.cy = (f.off_y + f.target * shot.target_y) / f.r;
    }
    if (shot.motion == "orbit") {   // circular Keplerian speed measured by the ZAMO
        double Om = 1.0 / (std::pow(f.r, 1.5) + a);
        double v = (Om - F.kp[6]) * F.kp[8] / F.kp[7];
        F.beta[2] = std::max(-0.98, std::min(0.98, v));
    }
    F.W = W;
    F.H = H;
    F.spp = spp;
    F.seed = seed;
    return F;
}

void camera_photon_host(double a, const bh_frame& F, double x, double y, double ph[9]) {
    V3 d = normalize(V3{F.F[0], F.F[1], F.F[2]} + V3{F.R[0], F.R[1], F.R[2]} * x + V3{F.U[0], F.U[1], F.U[2]} * y);
    V3 k = {-d.x, -d.y, -d.z}, beta = {F.beta[0], F.beta[1], F.beta[2]};
    double k0 = 1.0, b2 = dot(beta, beta);
    if (b2 > 0.0) {
        double b = std::sqrt(b2), g = 1.0 / std::sqrt(1.0 - b2);
        V3 bh = beta * (1.0 / b);
        double kpar = dot(k, bh);
        double k0z = g * (k0 + b * kpar);
        k = k + bh * ((g - 1.0) * kpar + g * b * k0);
        k0 = k0z;
    }
    const double* kp = F.kp;
    double L = kp[8] * k.z;
    double E = kp[7] * k0 + kp[6] * L;
    double pth = kp[9] * k.y;
    double Q = pth * pth + kp[2] * kp[2] * (L * L / (kp[1] * kp[1]) - a * a * E * E);
    ph[0] = L / E;
    ph[1] = Q / (E * E);
    ph[2] = E;
    ph[3] = kp[0];
    ph[4] = -(kp[10] * k.x / E);
    ph[5] = kp[2];
    ph[6] = kp[1] * (pth / E);
    ph[7] = F.ph;
    ph[8] = F.t;
}

bool critical_curve_point(double a, double theta_o, double rp, double* alpha, double* beta_plus) {
    double D = rp * rp - 2.0 * rp + a * a;
    double lam = a + rp / a * (rp - 2.0 * D / (rp - 1.0));
    double eta = rp * rp * rp / (a * a) * (4.0 * D / ((rp - 1.0) * (rp - 1.0)) - rp);
    double so = std::sin(theta_o), co = std::cos(theta_o);
    double b2 = eta + a * a * co * co - lam * lam * co * co / (so * so);
    if (b2 < 0.0) return false;
    *alpha = -lam / so;
    *beta_plus = std::sqrt(b2);
    return true;
}

bool find_critical_points(double a, const std::vector<bh_frame>& frames, const std::vector<double>& psis,
                          const FateFn& fates, std::vector<double>* rho) {
    const size_t q = frames.size();
    rho->assign(q, 0.0);
    if (q == 0) return true;
    std::vector<double> lo(q, 0.5), hi(q, 15.0), ph(9 * 2 * q);
    std::vector<int> fa;
    auto photon = [&](size_t i, double r, double* out) {
        double x = r * std::cos(psis[i]) / frames[i].r, y = r * std::sin(psis[i]) / frames[i].r;
        camera_photon_host(a, frames[i], x, y, out);
    };
    // bracket: captured at lo, escapes at hi
    for (size_t i = 0; i < q; ++i) {
        photon(i, lo[i], &ph[9 * (2 * i)]);
        photon(i, hi[i], &ph[9 * (2 * i + 1)]);
    }
    if (!fates(ph, &fa)) return false;
    std::vector<char> ok(q);
    for (size_t i = 0; i < q; ++i) ok[i] = fa[2 * i] == -1 && fa[2 * i + 1] == 1;
    ph.resize(9 * q);
    for (int it = 0; it < 64; ++it) {
        for (size_t i = 0; i < q; ++i) photon(i, 0.5 * (lo[i] + hi[i]), &ph[9 * i]);
        if (!fates(ph, &fa)) return false;
        for (size_t i = 0; i < q; ++i) {
            double mid = 0.5 * (lo[i] + hi[i]);
            if (fa[i] == 1) hi[i] = mid;
            else lo[i] = mid;
        }
    }
    for (size_t i = 0; i < q; ++i) (*rho)[i] = ok[i] ? 0.5 * (lo[i] + hi[i]) : 0.0;
    return true;
}

}  // namespace bh
