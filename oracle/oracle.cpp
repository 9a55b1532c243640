/* oracle.cpp - CPU FP64 restatement of bhrt's per-pixel Kerr tracer. TEST INFRASTRUCTURE ONLY
   (the checker for tests/, smoke() and bench.py's cpu_baseline; the product never links it).

   Every function follows the reference operation by operation (same expression order, so that with
   -ffp-contract=off it rounds identically), citing the reference file:line it restates. Paths are
   relative to the reference repository root. Out of scope and omitted: the thin disk
   (trace.cuh:106-128, emission.cuh:53-86), the sky texture (trace.cuh:35-43,154-160) and the
   inline fBm turbulence (emission.cuh:124-127). Turbulence: the reference's baked sheared noise
   (sheared.cpp) or a software-trilinear Cartesian density grid.

   ORACLE_BHM selects the elementary functions: 0 -> glibc (reference host build), 1 -> bh_math.h
   (the functions the GPU tracer uses). sqrt, division and fmin/fmax are IEEE on both sides. */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#if ORACLE_BHM
#include "../spacetime-optics_amd/csrc/bh_math.h"
#define M_POWM02(x) bhm::pow_m02(x)
#define M_EXP(x) bhm::exp(x)
#define M_LOG(x) bhm::log(x)
#define M_CBRT(x) bhm::cbrt(x)
#define M_EXPM1(x) bhm::expm1(x)
#define M_ATAN2(y, x) bhm::atan2(y, x)
#define M_ACOS(x) bhm::acos(x)
#define M_ASIN(x) bhm::asin(x)
static inline void M_SINCOS(double x, double* s, double* c) { bhm::sincos(x, s, c); }
#else
#define M_POWM02(x) std::pow(x, -0.2)
#define M_EXP(x) std::exp(x)
#define M_LOG(x) std::log(x)
#define M_CBRT(x) std::cbrt(x)
#define M_EXPM1(x) std::expm1(x)
#define M_ATAN2(y, x) std::atan2(y, x)
#define M_ACOS(x) std::acos(x)
#define M_ASIN(x) std::asin(x)
static inline void M_SINCOS(double x, double* s, double* c) {
    *s = std::sin(x);
    *c = std::cos(x);
}
#endif

namespace {

// blackholes/src/common.hpp:17-28
constexpr double C_C = 2.99792458e10, C_G = 6.67430e-8, C_KB = 1.380649e-16, C_H = 6.62607015e-27,
                 C_ME = 9.1093837e-28, C_MP = 1.67262192e-24, C_QE = 4.80320471e-10;
constexpr double PI = 3.14159265358979323846;

struct V3 {
    double x, y, z;
};
// blackholes/src/common.hpp:32-39
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return scl(a, 1.0 / length(a)); }
inline V3 v3(const double* p) { return {p[0], p[1], p[2]}; }

// blackholes/src/shot.cpp:43-46 (Rodrigues)
V3 rotate(V3 v, V3 axis, double ang) {
    double cs = std::cos(ang), sn = std::sin(ang);
    return add(add(scl(v, cs), scl(cross(axis, v), sn)), scl(axis, dot(axis, v) * (1 - cs)));
}

}  // namespace

extern "C" {

// blackholes/src/kerr.hpp:27-42
void or_kerr_point(double a, double r, double th, double k[11]) {
    double s = std::sin(th), c = std::cos(th);
    double a2 = a * a, r2 = r * r;
    double Sigma = r2 + a2 * c * c;
    double Delta = r2 - 2.0 * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * s * s;
    k[0] = r;
    k[1] = s;
    k[2] = c;
    k[3] = Sigma;
    k[4] = Delta;
    k[5] = A;
    k[6] = 2.0 * a * r / A;
    k[7] = std::sqrt(Sigma * Delta / A);
    k[8] = std::sqrt(A / Sigma) * s;
    k[9] = std::sqrt(Sigma);
    k[10] = std::sqrt(Sigma * Delta);
}

// blackholes/src/kerr.hpp:44
double or_horizon_radius(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

// blackholes/src/kerr.hpp:46-50
double or_isco_radius(double a) {
    double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    double z2 = std::sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

// blackholes/src/kerr.hpp:63-92; out = (lam, eta, E, y[6])
void or_camera_photon(double a, const double kp[11], double phi0, double t0, const double dd[3],
                      const double bb[3], double out[9]) {
    double k0 = 1.0;
    V3 d = v3(dd), beta = v3(bb);
    V3 k = neg(d);
    double b2 = dot(beta, beta);
    if (b2 > 0.0) {
        double b = std::sqrt(b2), g = 1.0 / std::sqrt(1.0 - b2);
        V3 bh = scl(beta, 1.0 / b);
        double kpar = dot(k, bh);
        double k0z = g * (k0 + b * kpar);
        k = add(k, scl(bh, (g - 1.0) * kpar + g * b * k0));
        k0 = k0z;
    }
    const double kr = kp[0], ks = kp[1], kc = kp[2], omega = kp[6], e_nu = kp[7], e_psi = kp[8];
    const double sqrtSigma = kp[9], sqrtSigmaDelta = kp[10];
    double L = e_psi * k.z;
    double E = e_nu * k0 + omega * L;
    double pth = sqrtSigma * k.y;
    double Q = pth * pth + kc * kc * (L * L / (ks * ks) - a * a * E * E);
    out[0] = L / E;
    out[1] = Q / (E * E);
    out[2] = E;
    double dr = sqrtSigmaDelta * k.x / E;
    double dth = pth / E;
    out[3] = kr;
    out[4] = -dr;
    out[5] = kc;
    out[6] = ks * dth;
    out[7] = phi0;
    out[8] = t0;
}

// blackholes/src/kerr.hpp:94-100
double or_radial_potential(double a, double lam, double eta, double r) {
    double P = r * r + a * a - a * lam, D = r * r - 2.0 * r + a * a;
    return P * P - D * (eta + (lam - a) * (lam - a));
}
double or_polar_potential(double a, double lam, double eta, double u) {
    return (1.0 - u * u) * (eta + a * a * u * u) - lam * lam * u * u;
}

// blackholes/src/kerr.hpp:103-116
void or_geodesic_rhs(double a, double lam, double eta, const double* y, double* d) {
    double r = y[0], u = y[2];
    double r2 = r * r, a2 = a * a;
    double Delta = r2 - 2.0 * r + a2;
    double P = r2 + a2 - a * lam;
    double K = eta + (lam - a) * (lam - a);
    double s2 = std::fmax(1.0 - u * u, 1e-300);
    d[0] = y[1];
    d[1] = 2.0 * r * P - (r - 1.0) * K;
    d[2] = y[3];
    d[3] = u * (a2 * (1.0 - 2.0 * u * u) - eta - lam * lam);
    d[4] = -(a * P / Delta + lam / s2 - a);
    d[5] = -((r2 + a2) * P / Delta + a * (lam - a * s2));
}

// blackholes/src/kerr.hpp:122-157 (Dormand-Prince 1980 tableau, FSAL, scaled max-norm over 0..4)
double or_dp54_step(double a, double lam, double eta, const double* y, const double* k1, double h, double tol,
                    double* yout, double* k7) {
    const double a21 = 1.0 / 5, a31 = 3.0 / 40, a32 = 9.0 / 40, a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
    const double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561, a54 = -212.0 / 729;
    const double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247, a64 = 49.0 / 176,
                 a65 = -5103.0 / 18656;
    const double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192, b5 = -2187.0 / 6784, b6 = 11.0 / 84;
    const double e1 = 71.0 / 57600, e3 = -71.0 / 16695, e4 = 71.0 / 1920, e5 = -17253.0 / 339200,
                 e6 = 22.0 / 525, e7 = -1.0 / 40;
    double k2[6], k3[6], k4[6], k5[6], k6[6], t[6];
    for (int i = 0; i < 6; ++i) t[i] = y[i] + h * a21 * k1[i];
    or_geodesic_rhs(a, lam, eta, t, k2);
    for (int i = 0; i < 6; ++i) t[i] = y[i] + h * (a31 * k1[i] + a32 * k2[i]);
    or_geodesic_rhs(a, lam, eta, t, k3);
    for (int i = 0; i < 6; ++i) t[i] = y[i] + h * (a41 * k1[i] + a42 * k2[i] + a43 * k3[i]);
    or_geodesic_rhs(a, lam, eta, t, k4);
    for (int i = 0; i < 6; ++i) t[i] = y[i] + h * (a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i]);
    or_geodesic_rhs(a, lam, eta, t, k5);
    for (int i = 0; i < 6; ++i)
        t[i] = y[i] + h * (a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i]);
    or_geodesic_rhs(a, lam, eta, t, k6);
    for (int i = 0; i < 6; ++i)
        yout[i] = y[i] + h * (b1 * k1[i] + b3 * k3[i] + b4 * k4[i] + b5 * k5[i] + b6 * k6[i]);
    or_geodesic_rhs(a, lam, eta, yout, k7);
    double sc[5];
    sc[0] = tol * (1e-6 + std::fmax(std::fabs(y[0]), std::fabs(yout[0])));
    sc[1] = tol * (1e-6 + std::fmax(std::fmax(std::fabs(y[1]), std::fabs(yout[1])), y[0] * y[0]));
    sc[2] = tol;
    sc[3] = tol * std::fmax(1.0, std::fmax(std::fabs(y[3]), std::fabs(yout[3])));
    sc[4] = tol * 4.0;
    double m = 0.0;
    for (int i = 0; i < 5; ++i) {
        double e = h * (e1 * k1[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * k7[i]);
        m = std::fmax(m, std::fabs(e) / sc[i]);
    }
    return m;
}

// blackholes/src/kerr.hpp:166-173
void or_project_radial(double a, double lam, double eta, double* y, double* k) {
    double r = y[0], R = or_radial_potential(a, lam, eta, r);
    if (R > 0.25 * r * r * r * r) {
        double v = std::sqrt(R);
        y[1] = y[1] < 0.0 ? -v : v;
        k[0] = y[1];
    }
}

// blackholes/src/kerr.hpp:175-178
double or_initial_step(const double* y) {
    double r = y[0];
    return 0.02 * std::fmin(r / (std::fabs(y[1]) + 1e-12), 1.0 / (std::fabs(y[3]) + 1e-3));
}

// blackholes/src/kerr.hpp:184-203
int or_vacuum_fate(double a, const double photon[9], double tol, double r_escape, int max_steps) {
    double rh = or_horizon_radius(a);
    const double lam = photon[0], eta = photon[1];
    double y[6], k1[6], yn[6], k7[6];
    for (int i = 0; i < 6; ++i) y[i] = photon[3 + i];
    or_geodesic_rhs(a, lam, eta, y, k1);
    double h = or_initial_step(y);
    for (int it = 0; it < max_steps; ++it) {
        double err = or_dp54_step(a, lam, eta, y, k1, h, tol, yn, k7);
        if (err > 1.0) {
            h *= std::fmax(0.2, 0.9 * M_POWM02(err));
            continue;
        }
        for (int i = 0; i < 6; ++i) {
            y[i] = yn[i];
            k1[i] = k7[i];
        }
        or_project_radial(a, lam, eta, y, k1);
        h *= std::fmin(5.0, std::fmax(0.2, 0.9 * M_POWM02(std::fmax(err, 1e-30))));
        if (y[0] < rh * 1.0001 + 1e-3) return -1;
        if (y[0] > r_escape && y[1] > 0.0) return 1;
    }
    return 0;
}

// blackholes/src/kerr.hpp:210-220
int or_critical_curve_point(double a, double theta_o, double rp, double* alpha, double* beta_plus) {
    double D = rp * rp - 2.0 * rp + a * a;
    double lam = a + rp / a * (rp - 2.0 * D / (rp - 1.0));
    double eta = rp * rp * rp / (a * a) * (4.0 * D / ((rp - 1.0) * (rp - 1.0)) - rp);
    double so = std::sin(theta_o), co = std::cos(theta_o);
    double b2 = eta + a * a * co * co - lam * lam * co * co / (so * so);
    if (b2 < 0.0) return 0;
    *alpha = -lam / so;
    *beta_plus = std::sqrt(b2);
    return 1;
}

// blackholes/src/shot.cpp:116-146 (angles in radians; fov in M when fov_in_m else radians)
void or_make_frame(double a, double r, double incl, double phi, double roll, double yaw, double pitch, double fov,
                   int fov_in_m, double zoom, double target, double target_x, double target_y, double off_x,
                   double off_y, double time_m, int orbit, int W, int H, uint32_t seed, or_frame* F) {
    std::memset(F, 0, sizeof(*F));
    F->r = r;
    F->th = incl;
    F->ph = phi;
    F->t = time_m;
    or_kerr_point(a, r, incl, F->kp);
    V3 Fw = {-1, 0, 0}, Up = {0, -1, 0}, Rt = {0, 0, 1};
    if (yaw != 0) {
        Fw = rotate(Fw, Up, -yaw);
        Rt = rotate(Rt, Up, -yaw);
    }
    if (pitch != 0) {
        Fw = rotate(Fw, Rt, pitch);
        Up = rotate(Up, Rt, pitch);
    }
    if (roll != 0) {
        Rt = rotate(Rt, Fw, roll);
        Up = rotate(Up, Fw, roll);
    }
    F->F[0] = Fw.x; F->F[1] = Fw.y; F->F[2] = Fw.z;
    F->R[0] = Rt.x; F->R[1] = Rt.y; F->R[2] = Rt.z;
    F->U[0] = Up.x; F->U[1] = Up.y; F->U[2] = Up.z;
    double width = fov_in_m ? fov / r : 2.0 * std::tan(0.5 * fov);
    F->pix = width / zoom / W;
    if (fov_in_m) {
        F->cx = (off_x + target * target_x) / r;
        F->cy = (off_y + target * target_y) / r;
    }
    if (orbit) {
        double Om = 1.0 / (std::pow(r, 1.5) + a);
        double v = (Om - F->kp[6]) * F->kp[8] / F->kp[7];
        F->beta[2] = std::fmax(-0.98, std::fmin(0.98, v));
    }
    F->W = W;
    F->H = H;
    F->spp = 1;
    F->seed = seed;
}

// blackholes/src/shot.cpp:148-164
double or_find_critical_point(double a, const or_frame* Fp, double psi) {
    const or_frame& F = *Fp;
    auto fate = [&](double rho) {
        double x = rho * std::cos(psi) / F.r, y = rho * std::sin(psi) / F.r;
        V3 d = normalize(add(add(v3(F.F), scl(v3(F.R), x)), scl(v3(F.U), y)));
        double dd[3] = {d.x, d.y, d.z}, ph[9];
        or_camera_photon(a, F.kp, F.ph, F.t, dd, F.beta, ph);
        return or_vacuum_fate(a, ph, 1e-12, 200.0, 2000000);
    };
    double lo = 0.5, hi = 15.0;
    if (fate(lo) != -1 || fate(hi) != 1) return 0.0;
    for (int it = 0; it < 64; ++it) {
        double mid = 0.5 * (lo + hi);
        if (fate(mid) == 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// blackholes/src/device_util.cuh:15-29
uint32_t or_pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static float rnd(uint32_t& s) {
    s = or_pcg(s);
    return (s >> 8) * (1.0f / 16777216.0f);
}
static float frac1(float x) { return x - std::floor(x); }
void or_r2_sample(uint32_t n, float r0, float r1, float* x, float* y) {
    const double a1 = 0.7548776662466927, a2 = 0.5698402909980532;
    *x = frac1(r0 + (float)std::fmod(n * a1, 1.0));
    *y = frac1(r1 + (float)std::fmod(n * a2, 1.0));
}
// blackholes/src/render.cu:70-74
void or_pixel_seed(uint32_t frame_seed, int p, float* r0, float* r1) {
    uint32_t h = or_pcg((uint32_t)p * 0x9E3779B9u ^ frame_seed);
    *r0 = rnd(h);
    *r1 = rnd(h);
}
// blackholes/src/render.cu:76-80 (jitter = 0: pixel centre, jx = jy = 0.5)
void or_sample_offsets(const or_frame* F, int px, int py, int s, int jitter, double* tx, double* ty) {
    float jx = 0.5f, jy = 0.5f;
    if (jitter) {
        float r0, r1;
        or_pixel_seed(F->seed, py * F->W + px, &r0, &r1);
        or_r2_sample((uint32_t)s, r0, r1, &jx, &jy);
    }
    *tx = (px + jx - 0.5 * F->W) * F->pix;
    *ty = (0.5 * F->H - (py + jy)) * F->pix;
}

// blackholes/src/device_util.cuh:74-80
float or_table_lerp(const float* t, int n, float x0, float dx, float x) {
    float f = (x - x0) / dx;
    f = std::fmin(std::fmax(f, 0.f), (float)(n - 2) + 0.9999f);
    int i = (int)f;
    f -= i;
    return t[i] + (t[i + 1] - t[i]) * f;
}

// blackholes/src/scene_build.cpp:159-170
void or_k2_table(int n, float* tab, float* l0f, float* dlf) {
    double l0 = std::log(0.02), l1 = std::log(1e4);
    for (int i = 0; i < n; ++i) {
        double th = std::exp(l0 + (l1 - l0) * i / (n - 1));
        tab[i] = (float)std::log(std::cyl_bessel_k(2.0, 1.0 / th));
    }
    *l0f = (float)l0;
    *dlf = (float)((l1 - l0) / (n - 1));
}

// Turbulent density factor: cell-centred trilinear interpolation of an n^3 grid over [-half, half]^3
// at the Cartesian point (R cos phi, R sin phi, r u), R = r sin(theta). Replaces the texture
// fetch of emission.cuh:112-123; the interpolation order below defines the value exactly.
static double grid_trilinear(const or_scene* S, const float* G, const double p[3]) {
    const int n = S->turb_n;
    int i0[3];
    double f[3];
    for (int d = 0; d < 3; ++d) {
        double g = (p[d] + S->turb_half) * S->turb_inv_dx - 0.5;
        g = std::fmin(std::fmax(g, 0.0), (double)(n - 1));
        int i = (int)g;
        if (i > n - 2) i = n - 2;
        i0[d] = i;
        f[d] = g - (double)i;
    }
    size_t nn = (size_t)n;
    size_t b = ((size_t)i0[2] * nn + (size_t)i0[1]) * nn + (size_t)i0[0];
    size_t sy = nn, sz = nn * nn;
    double v000 = G[b], v100 = G[b + 1], v010 = G[b + sy], v110 = G[b + sy + 1];
    double v001 = G[b + sz], v101 = G[b + sz + 1], v011 = G[b + sz + sy], v111 = G[b + sz + sy + 1];
    double c00 = v000 + (v100 - v000) * f[0];
    double c10 = v010 + (v110 - v010) * f[0];
    double c01 = v001 + (v101 - v001) * f[0];
    double c11 = v011 + (v111 - v011) * f[0];
    double c0 = c00 + (c10 - c00) * f[1];
    double c1 = c01 + (c11 - c01) * f[1];
    return c0 + (c1 - c0) * f[2];
}

// With time slices: linear in coordinate time t between the two slices around it, clamped to the
// first and last slice.
double or_turb_factor(const or_scene* S, double r, double u, double phi, double t) {
    double s = std::sqrt(std::fmax(1.0 - u * u, 1e-12));
    double R = r * s, sp, cp;
    M_SINCOS(phi, &sp, &cp);
    const double p[3] = {R * cp, R * sp, r * u};
    if (S->turb_nt <= 1) return grid_trilinear(S, S->turb_grid, p);
    double st = (t - S->turb_t0) * S->turb_inv_dt;
    st = std::fmin(std::fmax(st, 0.0), (double)(S->turb_nt - 1));
    int k = (int)st;
    if (k > S->turb_nt - 2) k = S->turb_nt - 2;
    const double f = st - (double)k;
    const size_t slice = (size_t)S->turb_n * S->turb_n * S->turb_n;
    const double A = grid_trilinear(S, S->turb_grid + (size_t)k * slice, p);
    const double B = grid_trilinear(S, S->turb_grid + (size_t)(k + 1) * slice, p);
    return A + (B - A) * f;
}

// blackholes/src/emission.cuh:20
static double keplerian_omega(const or_scene* S, double R) { return 1.0 / (R * std::sqrt(R) + S->a); }

// blackholes/src/emission.cuh:89-91
static double bessel_k2_inv(const or_scene* S, double theta) {
    return M_EXP((double)or_table_lerp(S->k2_lntab, S->k2_n, S->k2_l0, S->k2_dl, (float)M_LOG(theta)));
}

// blackholes/src/emission.cuh:99-175
void or_flow_step(const or_scene* S, double r, double u, double phi, double t, double lam, double Ecam, double h,
                  double* dI, double* dtau) {
    *dI = 0.0;
    *dtau = 0.0;
    const double a = S->a;
    double s2 = std::fmax(1.0 - u * u, 1e-12), s = std::sqrt(s2);
    double R = r * s, z = r * u;
    if (r > S->r_flow_max || R < 1e-3) return;
    double zh = z / (S->h_r * R);
    double vert = M_EXP(-0.5 * zh * zh);
    if (vert < 1e-7) return;
    double ne = S->n0 * M_EXP(-S->n_index * M_LOG(r)) * vert;
    if (S->flow_turb_amp > 0.f && S->turb_model == 1 && S->ts_slices)
        ne *= (double)or_sheared_factor(S, r, R, z, phi, t);
    else if (S->flow_turb_amp > 0.f && S->turb_grid && S->turb_n > 1)
        ne *= or_turb_factor(S, r, u, phi, t);
    ne *= 1.0 - M_EXP(-(r - S->rh) * (r - S->rh) / 0.25);
    if (ne < 1e-30) return;
    double Th = S->theta0 * M_EXP(-S->theta_index * M_LOG(r));
    double B = std::sqrt(8.0 * PI * ne * C_MP * C_C * C_C / (6.0 * r * S->beta_p));

    double r2 = r * r, a2 = a * a;
    double Sigma = r2 + a2 * u * u, Delta = r2 - 2.0 * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * s2;
    double g_tt = -(1.0 - 2.0 * r / Sigma), g_tph = -2.0 * a * r * s2 / Sigma, g_phph = A * s2 / Sigma;
    double omz = 2.0 * a * r / A;
    double Om = S->flow_kep * keplerian_omega(S, std::fmax(R, S->r_isco));
    double vmax = std::sqrt(Sigma * Delta / A) / std::sqrt(g_phph) * 0.999;
    Om = omz + std::fmax(-vmax, std::fmin(vmax, Om - omz));
    double norm = -(g_tt + 2.0 * g_tph * Om + g_phph * Om * Om);
    if (norm <= 0.0) return;
    double ut = 1.0 / std::sqrt(norm);
    double kdotu = ut * (1.0 - Om * lam);
    double nu_f = S->nu_obs * Ecam * kdotu;
    double g = S->nu_obs / nu_f;
    double u_ph = ut * (g_tph + g_phph * Om);
    double kb = lam - kdotu * u_ph;
    double bb = g_phph + u_ph * u_ph;
    double cosb = kb / (kdotu * std::sqrt(bb));
    double sinb = std::sqrt(std::fmax(1.0 - cosb * cosb, 1e-6));
    double nuc = C_QE * B / (2.0 * PI * C_ME * C_C);
    double nus = 2.0 / 9.0 * nuc * Th * Th * sinb;
    double X = nu_f / nus;
    double x3 = M_CBRT(X);
    double k2 = bessel_k2_inv(S, Th);
    double brk = std::sqrt(X) + 1.887748625363387 * std::sqrt(x3);
    double j = ne * 1.4142135623730951 * PI * C_QE * C_QE * nus / (3.0 * k2 * C_C) * brk * brk * M_EXP(-x3);
    double Te = Th * C_ME * C_C * C_C / C_KB;
    double hnu = C_H * nu_f / (C_KB * Te);
    double Bnu = 2.0 * C_H * nu_f * nu_f * nu_f / (C_C * C_C) / M_EXPM1(hnu);
    double alpha = j / Bnu;
    double dl = kdotu * Sigma * h * S->rg_cm;
    double tau = alpha * dl;
    *dtau = tau;
    *dI = g * g * g * Bnu * (-M_EXPM1(-tau));
}

// blackholes/src/trace.cuh:49-55
void or_escape_angles(const double* y, const double* k, double* th, double* ph) {
    double s = std::sqrt(std::fmax(1.0 - y[2] * y[2], 1e-30));
    double t0 = M_ACOS(std::fmax(-1.0, std::fmin(1.0, y[2])));
    double dthdr = (-y[3] / s) / y[1], dphdr = k[4] / y[1];
    *th = t0 + y[0] * dthdr;
    *ph = y[4] + y[0] * dphdr;
}

// blackholes/src/trace.cuh:25-33
void or_sky_uv(const double rot[9], double th, double ph, float* u, float* v) {
    double st, ct, sp, cp;
    M_SINCOS(th, &st, &ct);
    M_SINCOS(ph, &sp, &cp);
    const double d[3] = {st * cp, st * sp, ct};
    double g[3];
    for (int i = 0; i < 3; ++i) g[i] = rot[3 * i] * d[0] + rot[3 * i + 1] * d[1] + rot[3 * i + 2] * d[2];
    const double l = M_ATAN2(g[1], g[0]), b = M_ASIN(std::fmax(-1.0, std::fmin(1.0, g[2])));
    float uu = (float)(0.5 - l * (0.5 / PI));
    uu -= std::floor(uu);
    *u = uu;
    *v = (float)(0.5 - b / PI);
}

// The procedural sky of this path (the reference samples a star-map texture, trace.cuh:35-45)
float or_sky_checker(int nu, int nv, float scale, float u, float v) {
    int iu = (int)(u * (float)nu), iv = (int)(v * (float)nv);
    iu = std::min(std::max(iu, 0), nu - 1);
    iv = std::min(std::max(iv, 0), nv - 1);
    return ((iu + iv) & 1) ? scale : 0.1f * scale;
}

// blackholes/src/trace.cuh:62-165 without the disk branch (106-128); in visible mode an escaped ray
// takes the procedural checkerboard sky at its galactic (u, v) (the reference's star-map texture,
// 154-160); the intensity accumulates in FP64 (the reference sums floats into a float4).
void or_trace_ray(const or_scene* S, const or_frame* F, double x, double yv, or_ray_out* o) {
    std::memset(o, 0, sizeof(*o));
    const double a = S->a;
    V3 d = normalize(add(add(v3(F->F), scl(v3(F->R), F->cx + x)), scl(v3(F->U), F->cy + yv)));
    double dd[3] = {d.x, d.y, d.z}, ph[9];
    or_camera_photon(a, F->kp, F->ph, F->t, dd, F->beta, ph);
    double y[6];
    for (int i = 0; i < 6; ++i) y[i] = ph[3 + i];
    const double lam = ph[0], eta = ph[1], Ecam = ph[2];
    double k1[6], yn[6], k7[6];
    or_geodesic_rhs(a, lam, eta, y, k1);
    double h = or_initial_step(y);
    double acc[4] = {0, 0, 0, 0};
    double trans = 1.0;
    int ncross = 0, accepted = 0, flow_steps = 0, it = 0, fate = 0;
    double psi = 0.0;
    double sp = std::sqrt(std::fmax(1.0 - y[2] * y[2], 0.0));
    double s4, c4;
    M_SINCOS(y[4], &s4, &c4);
    V3 pprev = {sp * c4, sp * s4, y[2]};
    const bool radio = S->mode == 1;
    double th_inf = 0, ph_inf = 0;
    for (it = 0; it < S->max_steps; ++it) {
        double hc = h;
        if (radio && S->flow_on && y[0] < S->r_flow_max) {
            double s = std::sqrt(std::fmax(1.0 - y[2] * y[2], 0.0));
            double cap =
                S->step_frac * std::fmin(y[0] / (std::fabs(y[1]) + 1e-30), std::fmax(s, 0.05) / (std::fabs(y[3]) + 1e-30));
            hc = std::fmin(hc, cap);
        }
        double err = or_dp54_step(a, lam, eta, y, k1, hc, S->tol, yn, k7);
        if (!(err <= 1.0)) {
            h = hc * std::fmax(0.2, 0.9 * M_POWM02(std::fmax(err, 1e-30)));
            if (!(err == err)) h = hc * 0.1;
            continue;
        }
        ++accepted;
        if (radio) {
            double sn = std::sqrt(std::fmax(1.0 - yn[2] * yn[2], 0.0));
            double sv, cv;
            M_SINCOS(yn[4], &sv, &cv);
            V3 pn = {sn * cv, sn * sv, yn[2]};
            psi += M_ATAN2(length(cross(pprev, pn)), dot(pprev, pn));
            pprev = pn;
        }
        if ((y[2] > 0.0) != (yn[2] > 0.0)) ++ncross;
        if (radio && S->flow_on && yn[0] < S->r_flow_max) {
            double dI, dtau;
            ++flow_steps;
            or_flow_step(S, yn[0], yn[2], yn[4], yn[5], lam, Ecam, hc, &dI, &dtau);
            double c = trans * dI;
            int n = (int)(psi * (1.0 / PI));
            acc[0] += c;
            if (n == 1) acc[1] += c;
            else if (n == 2) acc[2] += c;
            else if (n >= 3) acc[3] += c;
            trans *= M_EXP(-dtau);
            if (trans < 1e-5) {
                fate = -1;
                break;
            }
        }
        for (int i = 0; i < 6; ++i) {
            y[i] = yn[i];
            k1[i] = k7[i];
        }
        or_project_radial(a, lam, eta, y, k1);
        h = hc * std::fmin(5.0, std::fmax(0.2, 0.9 * M_POWM02(std::fmax(err, 1e-30))));
        if (y[0] < S->r_stop) {
            fate = -1;
            break;
        }
        if (y[0] > S->r_escape && y[1] > 0.0) {
            fate = 1;
            or_escape_angles(y, k1, &th_inf, &ph_inf);
            if (!radio) {
                or_sky_uv(S->sky_rot, th_inf, ph_inf, &o->sky_u, &o->sky_v);
                if (S->sky_on) {
                    const double s = (double)or_sky_checker(S->sky_nu, S->sky_nv, S->sky_scale, o->sky_u, o->sky_v);
                    acc[0] += trans * s;
                    acc[1] += trans * s;
                    acc[2] += trans * s;
                }
            }
            break;
        }
    }
    for (int i = 0; i < 6; ++i) o->y[i] = y[i];
    o->th_inf = th_inf;
    o->ph_inf = ph_inf;
    for (int i = 0; i < 4; ++i) o->I[i] = acc[i];
    o->trans = trans;
    o->psi = psi;
    o->fate = fate;
    o->ncross = ncross;
    o->order = (int)(psi * (1.0 / PI));
    o->steps = it < S->max_steps ? it + 1 : S->max_steps;
    o->accepted = accepted;
    o->flow_steps = flow_steps;
}

int or_trace_pixels(const or_scene* S, const or_frame* F, const int* px, const int* py, int n, int spp, int jitter,
                    int nthreads, double* out_px, or_ray_out* out_rays) {
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads <= 0) nthreads = 1;
    std::atomic<int> next(0);
    const int chunk = 16;
    auto work = [&]() {
        for (;;) {
            int b = next.fetch_add(chunk);
            if (b >= n) break;
            int e = b + chunk < n ? b + chunk : n;
            for (int i = b; i < e; ++i) {
                double acc[4] = {0, 0, 0, 0};
                for (int s = 0; s < spp; ++s) {
                    double tx, ty;
                    or_sample_offsets(F, px[i], py[i], s, jitter, &tx, &ty);
                    or_ray_out o;
                    or_trace_ray(S, F, tx, ty, &o);
                    for (int c = 0; c < 4; ++c) acc[c] += o.I[c];
                    if (out_rays && spp == 1) out_rays[i] = o;
                }
                if (out_px)
                    for (int c = 0; c < 4; ++c) out_px[4 * (size_t)i + c] = acc[c] / spp;
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    return nthreads;
}

extern "C++" {
namespace {

// Runs f(i) for i in [0, n) on nthreads threads (dynamic scheduling in chunks of 16).
template <class Fn>
void parallel_for(int n, int nthreads, Fn f) {
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads <= 0) nthreads = 1;
    std::atomic<int> next(0);
    auto work = [&]() {
        for (;;) {
            int b = next.fetch_add(16);
            if (b >= n) break;
            int e = b + 16 < n ? b + 16 : n;
            for (int i = b; i < e; ++i) f(i);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

}  // namespace
}  // extern "C++"

// Sample s of pixel (px, py) (the ray render.cu:89-128 traces for it)
void or_trace_sample(const or_scene* S, const or_frame* F, int px, int py, int s, int jitter, or_ray_out* out) {
    double tx, ty;
    or_sample_offsets(F, px, py, s, jitter, &tx, &ty);
    or_trace_ray(S, F, tx, ty, out);
}

long long or_render_adaptive(const or_scene* S, const or_frame* F, int spp, int spp_max, float target_err,
                             int jitter, int nthreads, double* out_px, uint32_t* out_nsamp) {
    const int W = F->W, H = F->H, n = W * H, chunk = 8;
    const int lum = S->mode == 1 ? 0 : 1;
    std::vector<double> acc((size_t)n * 4, 0.0);
    std::vector<float> ys((size_t)n, 0.f), yq((size_t)n, 0.f);
    auto sample = [&](int p, int s, double* a, float* sy, float* sq) {
        double tx, ty;
        or_sample_offsets(F, p % W, p / W, s, jitter, &tx, &ty);
        or_ray_out o;
        or_trace_ray(S, F, tx, ty, &o);
        for (int c = 0; c < 4; ++c) a[c] += o.I[c];
        if (sy) {
            const float Y = (float)o.I[lum];
            *sy += Y;
            *sq += Y * Y;
        }
    };
    parallel_for(n, nthreads, [&](int p) {
        for (int s = 0; s < spp; ++s) sample(p, s, &acc[4 * (size_t)p], &ys[p], &yq[p]);
    });
    std::vector<uint32_t> chunks((size_t)n, 0u);
    const int extra_max = spp_max - spp > 0 ? spp_max - spp : 0;
    long long total = 0;
    if (extra_max > 0) {
        double fsum = 0.0;
        for (int b = 0; b < n; b += 256) {
            double part = 0.0;
            for (int p = b; p < n && p < b + 256; ++p) part += (double)ys[p];
            fsum += part;
        }
        const float yref = (float)(0.03 * fsum / ((double)n * spp) + 1e-37);
        std::vector<uint32_t> extra((size_t)n);
        for (int p = 0; p < n; ++p) {   // render.cu:130-139
            const float m = (float)spp;
            const float mean = ys[p] / m, var = std::fmax(yq[p] / m - mean * mean, 0.f);
            const float se = std::sqrt(var / m), den = target_err * (mean + yref);
            const float want = m * (se / den) * (se / den);
            extra[p] = (uint32_t)std::fmin(std::fmax(std::ceil(want) - m, 0.f), (float)extra_max);
        }
        for (int y = 0; y < H; ++y)   // render.cu:141-150
            for (int x = 0; x < W; ++x) {
                uint32_t mx = 0;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int yy = std::min(std::max(y + dy, 0), H - 1), xx = std::min(std::max(x + dx, 0), W - 1);
                        mx = std::max(mx, extra[(size_t)yy * W + xx]);
                    }
                chunks[(size_t)y * W + x] = (std::min(mx, (uint32_t)extra_max) + chunk - 1) / chunk;
            }
        for (int p = 0; p < n; ++p) total += chunks[p];
        parallel_for(n, nthreads, [&](int p) {
            for (uint32_t k = 0; k < chunks[p] * chunk; ++k) sample(p, spp + (int)k, &acc[4 * (size_t)p], nullptr, nullptr);
        });
    }
    for (int p = 0; p < n; ++p) {
        const uint32_t m = (uint32_t)spp + chunks[p] * chunk;
        for (int c = 0; c < 4; ++c) out_px[4 * (size_t)p + c] = acc[4 * (size_t)p + c] / m;
        if (out_nsamp) out_nsamp[p] = m;
    }
    return (long long)n * spp + total * chunk;
}

void or_math_eval(int fn, const double* x, const double* y, double* out, int n) {
    for (int i = 0; i < n; ++i) {
        double a = x[i], b = y[i], s, c, r;
        switch (fn) {
            case 0: r = M_EXP(a); break;
            case 1: r = M_LOG(a); break;
            case 2: r = M_EXPM1(a); break;
            case 3: r = M_CBRT(a); break;
            case 4: r = M_POWM02(a); break;
            case 5: M_SINCOS(a, &s, &c); r = s; break;
            case 6: M_SINCOS(a, &s, &c); r = c; break;
            case 7: r = M_ATAN2(a, b); break;
            case 8: r = M_ACOS(a); break;
            case 9: r = std::sqrt(a); break;
            case 11: r = M_ASIN(a); break;
            default: r = a / b; break;
        }
        out[i] = r;
    }
}

// blackholes/src/main.cpp:323-372
void or_selftest(double rows[8][6], double* drift, int* drift_steps) {
    const double spins[4] = {1e-6, 0.5, 0.94, 0.99}, incls[2] = {90.0, 17.0};
    int row = 0;
    for (double a : spins) {
        for (double incl_deg : incls) {
            double incl = incl_deg * PI / 180;
            double amin = 1e9, amax = -1e9;
            double rp0 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(-a))),
                   rp1 = 2.0 * (1.0 + std::cos(2.0 / 3.0 * std::acos(a)));
            for (int i = 0; i <= 200000; ++i) {
                double rp = rp0 + (rp1 - rp0) * i / 200000.0, al, be;
                if (or_critical_curve_point(a, incl, rp, &al, &be)) {
                    amin = std::fmin(amin, al);
                    amax = std::fmax(amax, al);
                }
            }
            or_frame F;
            or_make_frame(a, 1e7, incl, 0, 0, 0, 0, 30.0, 1, 1.0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, &F);
            double xp = or_find_critical_point(a, &F, 0.0);
            double xm = or_find_critical_point(a, &F, PI);
            rows[row][0] = a;
            rows[row][1] = incl_deg;
            rows[row][2] = -xm;
            rows[row][3] = xp;
            rows[row][4] = amin;
            rows[row][5] = amax;
            ++row;
        }
    }
    double a = 0.94, kp[11], ph[9];
    or_kerr_point(a, 50.0, 1.2, kp);
    V3 dv = normalize(V3{-1, 0.13, 0.11});
    double dd[3] = {dv.x, dv.y, dv.z}, b0[3] = {0, 0, 0};
    or_camera_photon(a, kp, 0, 0, dd, b0, ph);
    double y[6], k1[6], yn[6], k7[6];
    for (int i = 0; i < 6; ++i) y[i] = ph[3 + i];
    const double lam = ph[0], eta = ph[1];
    or_geodesic_rhs(a, lam, eta, y, k1);
    double h = or_initial_step(y), worst = 0;
    int steps = 0;
    for (int it = 0; it < 200000 && steps < 3000; ++it) {
        double err = or_dp54_step(a, lam, eta, y, k1, h, 1e-10, yn, k7);
        if (err > 1) {
            h *= std::fmax(0.2, 0.9 * M_POWM02(err));
            continue;
        }
        for (int i = 0; i < 6; ++i) {
            y[i] = yn[i];
            k1[i] = k7[i];
        }
        h *= std::fmin(5.0, std::fmax(0.2, 0.9 * M_POWM02(std::fmax(err, 1e-30))));
        ++steps;
        double R = or_radial_potential(a, lam, eta, y[0]);
        worst = std::fmax(worst, std::fabs(R - y[1] * y[1]) / std::pow(y[0], 4.0));
        if (y[0] < or_horizon_radius(a) + 0.01 || y[0] > 1e4) break;
    }
    *drift = worst;
    *drift_steps = steps;
}

}  // extern "C"
