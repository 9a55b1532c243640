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

// Reads "t0:v0, t1:v1, ..." (seconds : value) into a track; scale-like quantities (log) interpolate
// ln v. Empty entries are skipped. False, with the reason, when an entry is not two numbers joined by
// ':' or when there is no entry at all.
bool read_track(const std::string& text, bool log, Track* tr, std::string* why) {
    tr->t.clear();
    tr->v.clear();
    tr->log = log;
    for (size_t pos = 0; pos <= text.size();) {
        size_t comma = text.find(',', pos);
        if (comma == std::string::npos) comma = text.size();
        const std::string entry = trim(text.substr(pos, comma - pos));
        pos = comma + 1;
        if (entry.empty()) continue;
        const size_t colon = entry.find(':');
        if (colon == std::string::npos) {
            *why = "'" + entry + "' is not time:value";
            return false;
        }
        tr->t.push_back(std::strtod(entry.c_str(), nullptr));   // stops at the ':'
        const double v = std::strtod(entry.c_str() + colon + 1, nullptr);
        tr->v.push_back(log ? std::log(v) : v);
    }
    if (tr->t.empty()) *why = "no time:value entries";
    return !tr->t.empty();
}

// One timeline quantity of a shot, resolved once: its [track] curve when "track.<key>" holds a valid
// one (an invalid one is reported and ignored), else the key's constant (or the default).
struct Channel {
    bool tracked = false;
    double constant = 0.0;
    Track curve;
    double at(double time) const { return tracked ? curve.value(time) : constant; }
};

Channel resolve(const Config& c, const std::string& key, double dflt, bool log) {
    Channel ch;
    const std::string tk = "track." + key;
    if (c.has(tk)) {
        std::string why;
        ch.tracked = read_track(c.str(tk), log, &ch.curve, &why);
        if (!ch.tracked) std::fprintf(stderr, "warning: [track] %s: %s; using the constant\n", key.c_str(), why.c_str());
    }
    if (!ch.tracked) ch.constant = c.num(key, dflt);
    return ch;
}

// The per-frame quantities that are plain channels: key, default, interpolated in log space, given
// in degrees (stored in radians), destination (shot.cpp:78-99 reads the same keys and defaults).
struct ChannelSpec {
    const char* key;
    double dflt;
    bool log, degrees;
    double ShotFrame::*field;
};
const ChannelSpec TIMELINE[] = {
    {"camera.r", 1e6, true, false, &ShotFrame::r},
    {"camera.incl", 80, false, true, &ShotFrame::incl},
    {"camera.phi", 0, false, true, &ShotFrame::phi},
    {"camera.roll", 0, false, true, &ShotFrame::roll},
    {"camera.yaw", 0, false, true, &ShotFrame::yaw},
    {"camera.pitch", 0, false, true, &ShotFrame::pitch},
    {"camera.zoom", 1, true, false, &ShotFrame::zoom},
    {"camera.target", 0, false, false, &ShotFrame::target},
    {"camera.offset_x_m", 0, false, false, &ShotFrame::off_x},
    {"camera.offset_y_m", 0, false, false, &ShotFrame::off_y},
    {"time.t_m", 0, false, false, &ShotFrame::time_m},
    // post-processing timeline: beam blur, order tint, direct-image removal, exposure offset, HUD opacity
    {"post.blur", 0, false, false, &ShotFrame::blur},
    {"post.overlay", 0, false, false, &ShotFrame::overlay},
    {"post.hide_direct", 0, false, false, &ShotFrame::hide_direct},
    {"post.exposure_ev", 0, false, false, &ShotFrame::exposure_ev},
    {"hud.alpha", 1, false, false, &ShotFrame::hud},
};

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

// The shot's frame list (shot.cpp:73-114): frame i at t = i / fps, every quantity sampled from its
// channel; the field of view in M (camera.fov_m) or degrees (camera.fov_deg); the fade is post.fade
// times linear ramps over shot.fade_in and shot.fade_out seconds at the two ends.
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
    const int by_time = std::max(1, (int)std::lround(shot->duration * shot->fps));
    const int count = c.has("shot.frames") ? c.integer("shot.frames", by_time) : by_time;
    const double deg = PI / 180;
    constexpr size_t NCH = sizeof(TIMELINE) / sizeof(TIMELINE[0]);
    std::vector<Channel> chans;
    chans.reserve(NCH);
    for (const ChannelSpec& sp : TIMELINE) chans.push_back(resolve(c, sp.key, sp.dflt, sp.log));
    const bool fov_in_m = c.has("camera.fov_m") || c.has("track.camera.fov_m");
    const Channel fov = fov_in_m ? resolve(c, "camera.fov_m", 30, true) : resolve(c, "camera.fov_deg", 40, true);
    const Channel fade = resolve(c, "post.fade", 1, false);
    const double fade_in = c.num("shot.fade_in", 0.0), fade_out = c.num("shot.fade_out", 0.0);

    shot->frames.assign((size_t)std::max(count, 0), ShotFrame{});
    for (int i = 0; i < count; ++i) {
        ShotFrame& f = shot->frames[(size_t)i];
        f.index = i;
        f.t = i / shot->fps;
        for (size_t k = 0; k < NCH; ++k) {
            const double v = chans[k].at(f.t);
            f.*(TIMELINE[k].field) = TIMELINE[k].degrees ? v * deg : v;
        }
        f.fov_in_m = fov_in_m;
        f.fov = fov_in_m ? fov.at(f.t) : fov.at(f.t) * deg;
        double level = fade.at(f.t);
        if (fade_in > 0) level *= std::min(1.0, f.t / fade_in);
        if (fade_out > 0) level *= std::min(1.0, (shot->duration - f.t) / fade_out);
        f.fade = std::max(0.0, level);
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
