/* shot.h - camera timeline and frame construction (blackholes/src/shot.hpp:1-50 of the reference):
   parameters constant in their section or animated in [track] as "time:value, ..." (seconds),
   monotone-cubic interpolation, in log space for scale-like quantities. */
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "../../../include/bhrt.h"
#include "config.h"

namespace bh {

struct Track {
    std::vector<double> t, v;
    bool log = false;
    double value(double time) const;
};

struct ShotFrame {
    int index = 0;
    double t = 0;
    double r = 0, incl = 0, phi = 0;   // M, rad, rad
    double roll = 0, yaw = 0, pitch = 0;
    double fov = 0;                    // M (fov_in_m) or radians
    bool fov_in_m = false;
    double zoom = 1, target = 0, off_x = 0, off_y = 0, time_m = 0;
    double blur = 0, overlay = 0, hide_direct = 0, exposure_ev = 0, fade = 1, hud = 1;   // post timeline
};

struct Shot {
    std::string name, type, motion;
    double duration = 5, fps = 30;
    double target_x = 0, target_y = 0;
    std::vector<ShotFrame> frames;
};

bool build_shot(const Config& c, Shot* shot, std::string* err);

// make_frame (shot.cpp:116-146): camera of a frame -> bh_frame (W, H, spp, seed filled in)
bh_frame make_frame(const Shot& shot, const ShotFrame& f, double a, int W, int H, int spp, uint32_t seed);

// camera_photon (kerr.hpp:63-92) on the host: ph = (lam, eta, E, y[6]) of image-plane offset (x, y)
void camera_photon_host(double a, const bh_frame& F, double x, double y, double ph[9]);

// critical_curve_point (kerr.hpp:210-220): analytic Bardeen shadow edge
bool critical_curve_point(double a, double theta_o, double rp, double* alpha, double* beta_plus);

// Vacuum fates of photon batches (device); returns false on error.
using FateFn = std::function<bool(const std::vector<double>& photons, std::vector<int>* fates)>;

/* find_critical_point (shot.cpp:148-164) for several (frame, psi) queries at once: 64 bisection
   steps on rho in [0.5, 15] M along image direction psi; every step evaluates the midpoints of all
   queries in one device batch. rho = 0 where the bracket does not hold. */
bool find_critical_points(double a, const std::vector<bh_frame>& frames, const std::vector<double>& psis,
                          const FateFn& fates, std::vector<double>* rho);

}  // namespace bh
