/* trace_fp64.h - plain FP64 per-ray Kerr tracer for gfx950 (device code, included by bhrt.hip).

   Parity contract: with -ffp-contract=off every value below rounds exactly like the reference
   integrator's host build (blackholes/src/kerr.hpp, trace.cuh, emission.cuh), except that the
   elementary functions come from bh_math.h; the CPU oracle built on bh_math.h reproduces the
   results bit for bit.

   MI355X layout of the work:
   - One ray per lane, wave64; the integrator state lives in VGPRs for the whole trace.
   - Dormand-Prince 5(4) is evaluated incrementally: after each stage derivative k_i the partial
     sums of every later stage, of the 5th-order solution and of the error estimate are updated in
     the reference's left-to-right association order, so the live set is partial sums instead of
     seven stored derivative vectors.
   - The phi and t components never enter the right-hand side (kerr.hpp:103-116), so their stage
     states are never formed; only their 5th-order sums (and phi's error term) are kept.
   - d[0] = y[1] and d[2] = y[3]: the r and u components of each stage derivative are the stage
     state's r' and u' and are not stored separately. */
#pragma once
#include "bh_math.h"

namespace bhd {

// blackholes/src/common.hpp:17-28
constexpr double C_C = 2.99792458e10, C_KB = 1.380649e-16, C_H = 6.62607015e-27, C_ME = 9.1093837e-28,
                 C_MP = 1.67262192e-24, C_QE = 4.80320471e-10;
constexpr double PI = 3.14159265358979323846;

// The tracers' RADIO template argument: no flow (thermal disk), or the radio flow with the grid
// turbulence or with the sheared turbulence. The turbulence model is a compile-time choice so a
// kernel carries the lookup code of one model only (registers and instruction cache).
constexpr int RADIO_OFF = 0, RADIO_GRID = 1, RADIO_SHEARED = 2;
constexpr int MODE_VISIBLE = 0;   // bh_scene.mode: vacuum geodesics and the celestial sphere

// Dormand-Prince 1980 tableau (kerr.hpp:124-129): the same constant expressions, folded by the compiler.
constexpr double a21 = 1.0 / 5, a31 = 3.0 / 40, a32 = 9.0 / 40, a41 = 44.0 / 45, a42 = -56.0 / 15, a43 = 32.0 / 9;
constexpr double a51 = 19372.0 / 6561, a52 = -25360.0 / 2187, a53 = 64448.0 / 6561, a54 = -212.0 / 729;
constexpr double a61 = 9017.0 / 3168, a62 = -355.0 / 33, a63 = 46732.0 / 5247, a64 = 49.0 / 176, a65 = -5103.0 / 18656;
constexpr double b1 = 35.0 / 384, b3 = 500.0 / 1113, b4 = 125.0 / 192, b5 = -2187.0 / 6784, b6 = 11.0 / 84;
constexpr double e1 = 71.0 / 57600, e3 = -71.0 / 16695, e4 = 71.0 / 1920, e5 = -17253.0 / 339200, e6 = 22.0 / 525,
                 e7 = -1.0 / 40;

// Baked sheared-turbulence slices (turb_sheared.h): nt slices of nz x nlr x nphi floats, phi fastest,
// with the slice coordinates and the camera they were baked for.
struct ShearedView {
    const float* data;
    int nphi, nlr, nz, nt;
    float lr0, lr1, zmax, d0, dd;
    double cam_t, cam_r;
};

// Uniform scene and frame parameters: passed by value as the kernel argument, so they sit in SGPRs.
struct TraceParams {
    // scene (bh_scene)
    double a, a2, rh, r_isco, tol, tol4, r_escape, r_stop, step_frac;
    int max_steps, mode, flow_on, turb_on;
    double n0, n_index, theta_index, theta0, h_r, beta_p, r_flow_max, nu_obs, rg_cm, flow_kep;
    const float* __restrict__ k2;
    int k2_n;
    float k2_l0, k2_dl;
    const float* __restrict__ turb;
    int turb_n, turb_nt;
    double turb_half, turb_inv_dx, turb_t0, turb_inv_dt;
    // sheared turbulence (RADIO_SHEARED kernels): the slices and their coordinates, in device memory
    // so the kernel arguments stay small
    const struct ShearedView* __restrict__ ts;
    // frame (bh_frame)
    double ph0, t0;
    double kp[11];
    double F[3], R[3], U[3], beta[3];
    double cx, cy, pix;
    int W, H;
    uint32_t seed;
    // visible mode: checkerboard celestial sphere (sky_nu x sky_nv squares of galactic u, v)
    int sky_on, sky_nu, sky_nv;
    float sky_scale;
    double sky_rot[9];
    // launch
    int x0, y0, w, h, spp, jitter;
};

struct RayResult {
    double y[6];
    double th_inf, ph_inf;
    double I[4];
    double trans, psi;
    int fate, ncross, order, steps, accepted, flow_steps;
    float sky_u, sky_v;
};

struct RayConst {
    double lam, eta, E, K, lam2, alam;
};

// blackholes/src/device_util.cuh:15-29, render.cu:70-80
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
// the reference's sheared turbulence (bake and lookup), in this namespace
#include "turb_sheared.h"

__device__ __forceinline__ void sample_offsets(const TraceParams& P, int px, int py, int s, double* tx, double* ty) {
    float jx = 0.5f, jy = 0.5f;
    if (P.jitter) {
        uint32_t hs = pcg((uint32_t)(py * P.W + px) * 0x9E3779B9u ^ P.seed);
        hs = pcg(hs);
        float r0 = (hs >> 8) * (1.0f / 16777216.0f);
        hs = pcg(hs);
        float r1 = (hs >> 8) * (1.0f / 16777216.0f);
        // fmod(n a, 1) for n a >= 0 is n a - floor(n a), exactly
        double q1 = (double)(uint32_t)s * 0.7548776662466927, q2 = (double)(uint32_t)s * 0.5698402909980532;
        float fx = r0 + (float)(q1 - floor(q1)), fy = r1 + (float)(q2 - floor(q2));
        jx = fx - floorf(fx);
        jy = fy - floorf(fy);
    }
    *tx = (px + jx - 0.5 * P.W) * P.pix;
    *ty = (0.5 * P.H - (py + jy)) * P.pix;
}

// blackholes/src/kerr.hpp:103-116 (components 1, 3, 4, 5; d0 = y1 and d2 = y3)
struct Rhs {
    double d1, d3, d4, d5;
};
__device__ __forceinline__ Rhs rhs(const TraceParams& P, const RayConst& c, double r, double u) {
    double r2 = r * r;
    double Delta = r2 - 2.0 * r + P.a2;
    double Pp = r2 + P.a2 - c.alam;
    double s2 = fmax(1.0 - u * u, 1e-300);
    Rhs o;
    o.d1 = 2.0 * r * Pp - (r - 1.0) * c.K;
    o.d3 = u * (P.a2 * (1.0 - 2.0 * u * u) - c.eta - c.lam2);
    o.d4 = -(P.a * Pp / Delta + c.lam / s2 - P.a);
    o.d5 = -((r2 + P.a2) * Pp / Delta + P.a * (c.lam - P.a * s2));
    return o;
}

// Integrator state: y and the FSAL derivative k1 (k1[0] == y[1] and k1[2] == y[3] always hold).
struct State {
    double y0, y1, y2, y3, y4, y5;
    double k1, k3, k4, k5;   // k1[1], k1[3], k1[4], k1[5]
};

// blackholes/src/kerr.hpp:122-157. Returns the scaled error; on return `o` holds (yout, k7).
__device__ __forceinline__ double dp54_step(const TraceParams& P, const RayConst& c, const State& s, double h,
                                            State& o) {
    // partial sums of stages 3..6, of the 5th-order increment (O) and of the error (E), per component
    // (0: r, 1: r', 2: u, 3: u'); k_i component 0 is stage i's r', component 2 is stage i's u'.
    const double ha21 = h * a21;
    double q0 = s.y1, q1 = s.k1, q2 = s.y3, q3 = s.k3;   // k1 components 0..3
    double S3_0 = a31 * q0, S3_1 = a31 * q1, S3_2 = a31 * q2, S3_3 = a31 * q3;
    double S4_0 = a41 * q0, S4_1 = a41 * q1, S4_2 = a41 * q2, S4_3 = a41 * q3;
    double S5_0 = a51 * q0, S5_1 = a51 * q1, S5_2 = a51 * q2, S5_3 = a51 * q3;
    double S6_0 = a61 * q0, S6_1 = a61 * q1, S6_2 = a61 * q2, S6_3 = a61 * q3;
    double O0 = b1 * q0, O1 = b1 * q1, O2 = b1 * q2, O3 = b1 * q3, O4 = b1 * s.k4, O5 = b1 * s.k5;
    double E0 = e1 * q0, E1 = e1 * q1, E2 = e1 * q2, E3 = e1 * q3, E4 = e1 * s.k4;
    // stage 2
    double t0 = s.y0 + ha21 * q0, t1 = s.y1 + ha21 * q1, t2 = s.y2 + ha21 * q2, t3 = s.y3 + ha21 * q3;
    Rhs k = rhs(P, c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S3_0 = S3_0 + a32 * q0; S3_1 = S3_1 + a32 * q1; S3_2 = S3_2 + a32 * q2; S3_3 = S3_3 + a32 * q3;
    S4_0 = S4_0 + a42 * q0; S4_1 = S4_1 + a42 * q1; S4_2 = S4_2 + a42 * q2; S4_3 = S4_3 + a42 * q3;
    S5_0 = S5_0 + a52 * q0; S5_1 = S5_1 + a52 * q1; S5_2 = S5_2 + a52 * q2; S5_3 = S5_3 + a52 * q3;
    S6_0 = S6_0 + a62 * q0; S6_1 = S6_1 + a62 * q1; S6_2 = S6_2 + a62 * q2; S6_3 = S6_3 + a62 * q3;
    // stage 3
    t0 = s.y0 + h * S3_0; t1 = s.y1 + h * S3_1; t2 = s.y2 + h * S3_2; t3 = s.y3 + h * S3_3;
    k = rhs(P, c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S4_0 = S4_0 + a43 * q0; S4_1 = S4_1 + a43 * q1; S4_2 = S4_2 + a43 * q2; S4_3 = S4_3 + a43 * q3;
    S5_0 = S5_0 + a53 * q0; S5_1 = S5_1 + a53 * q1; S5_2 = S5_2 + a53 * q2; S5_3 = S5_3 + a53 * q3;
    S6_0 = S6_0 + a63 * q0; S6_1 = S6_1 + a63 * q1; S6_2 = S6_2 + a63 * q2; S6_3 = S6_3 + a63 * q3;
    O0 = O0 + b3 * q0; O1 = O1 + b3 * q1; O2 = O2 + b3 * q2; O3 = O3 + b3 * q3; O4 = O4 + b3 * k.d4;
    O5 = O5 + b3 * k.d5;
    E0 = E0 + e3 * q0; E1 = E1 + e3 * q1; E2 = E2 + e3 * q2; E3 = E3 + e3 * q3; E4 = E4 + e3 * k.d4;
    // stage 4
    t0 = s.y0 + h * S4_0; t1 = s.y1 + h * S4_1; t2 = s.y2 + h * S4_2; t3 = s.y3 + h * S4_3;
    k = rhs(P, c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S5_0 = S5_0 + a54 * q0; S5_1 = S5_1 + a54 * q1; S5_2 = S5_2 + a54 * q2; S5_3 = S5_3 + a54 * q3;
    S6_0 = S6_0 + a64 * q0; S6_1 = S6_1 + a64 * q1; S6_2 = S6_2 + a64 * q2; S6_3 = S6_3 + a64 * q3;
    O0 = O0 + b4 * q0; O1 = O1 + b4 * q1; O2 = O2 + b4 * q2; O3 = O3 + b4 * q3; O4 = O4 + b4 * k.d4;
    O5 = O5 + b4 * k.d5;
    E0 = E0 + e4 * q0; E1 = E1 + e4 * q1; E2 = E2 + e4 * q2; E3 = E3 + e4 * q3; E4 = E4 + e4 * k.d4;
    // stage 5
    t0 = s.y0 + h * S5_0; t1 = s.y1 + h * S5_1; t2 = s.y2 + h * S5_2; t3 = s.y3 + h * S5_3;
    k = rhs(P, c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S6_0 = S6_0 + a65 * q0; S6_1 = S6_1 + a65 * q1; S6_2 = S6_2 + a65 * q2; S6_3 = S6_3 + a65 * q3;
    O0 = O0 + b5 * q0; O1 = O1 + b5 * q1; O2 = O2 + b5 * q2; O3 = O3 + b5 * q3; O4 = O4 + b5 * k.d4;
    O5 = O5 + b5 * k.d5;
    E0 = E0 + e5 * q0; E1 = E1 + e5 * q1; E2 = E2 + e5 * q2; E3 = E3 + e5 * q3; E4 = E4 + e5 * k.d4;
    // stage 6
    t0 = s.y0 + h * S6_0; t1 = s.y1 + h * S6_1; t2 = s.y2 + h * S6_2; t3 = s.y3 + h * S6_3;
    k = rhs(P, c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    O0 = O0 + b6 * q0; O1 = O1 + b6 * q1; O2 = O2 + b6 * q2; O3 = O3 + b6 * q3; O4 = O4 + b6 * k.d4;
    O5 = O5 + b6 * k.d5;
    E0 = E0 + e6 * q0; E1 = E1 + e6 * q1; E2 = E2 + e6 * q2; E3 = E3 + e6 * q3; E4 = E4 + e6 * k.d4;
    // 5th-order solution and its derivative (FSAL)
    o.y0 = s.y0 + h * O0; o.y1 = s.y1 + h * O1; o.y2 = s.y2 + h * O2; o.y3 = s.y3 + h * O3;
    o.y4 = s.y4 + h * O4; o.y5 = s.y5 + h * O5;
    k = rhs(P, c, o.y0, o.y2);
    o.k1 = k.d1; o.k3 = k.d3; o.k4 = k.d4; o.k5 = k.d5;
    E0 = E0 + e7 * o.y1; E1 = E1 + e7 * k.d1; E2 = E2 + e7 * o.y3; E3 = E3 + e7 * k.d3; E4 = E4 + e7 * k.d4;
    // scaled max-norm over r, r', u, u', phi (kerr.hpp:145-155)
    double sc0 = P.tol * (1e-6 + fmax(fabs(s.y0), fabs(o.y0)));
    double sc1 = P.tol * (1e-6 + fmax(fmax(fabs(s.y1), fabs(o.y1)), s.y0 * s.y0));
    double sc3 = P.tol * fmax(1.0, fmax(fabs(s.y3), fabs(o.y3)));
    double m = 0.0;
    m = fmax(m, fabs(h * E0) / sc0);
    m = fmax(m, fabs(h * E1) / sc1);
    m = fmax(m, fabs(h * E2) / P.tol);
    m = fmax(m, fabs(h * E3) / sc3);
    m = fmax(m, fabs(h * E4) / P.tol4);
    return m;
}

// blackholes/src/kerr.hpp:166-173
__device__ __forceinline__ void project_radial(const TraceParams& P, const RayConst& c, State& s) {
    double r = s.y0;
    double Pp = r * r + P.a2 - c.alam, D = r * r - 2.0 * r + P.a2;
    double R = Pp * Pp - D * c.K;
    if (R > 0.25 * r * r * r * r) {
        double v = sqrt(R);
        s.y1 = s.y1 < 0.0 ? -v : v;
    }
}

// blackholes/src/device_util.cuh:74-80
__device__ __forceinline__ float table_lerp(const float* __restrict__ t, int n, float x0, float dx, float x) {
    float f = (x - x0) / dx;
    f = fminf(fmaxf(f, 0.f), (float)(n - 2) + 0.9999f);
    int i = (int)f;
    f -= i;
    float t0 = t[i], t1 = t[i + 1];
    return t0 + (t1 - t0) * f;
}

// Trilinear density factor on the Cartesian grid (slice pointer g), same order as the oracle.
__device__ __forceinline__ double grid_trilinear(const TraceParams& P, const float* __restrict__ G, double px,
                                                 double py, double pz) {
    const int n = P.turb_n;
    double gx = (px + P.turb_half) * P.turb_inv_dx - 0.5;
    double gy = (py + P.turb_half) * P.turb_inv_dx - 0.5;
    double gz = (pz + P.turb_half) * P.turb_inv_dx - 0.5;
    const double hi = (double)(n - 1);
    gx = fmin(fmax(gx, 0.0), hi);
    gy = fmin(fmax(gy, 0.0), hi);
    gz = fmin(fmax(gz, 0.0), hi);
    int ix = (int)gx, iy = (int)gy, iz = (int)gz;
    ix = ix > n - 2 ? n - 2 : ix;
    iy = iy > n - 2 ? n - 2 : iy;
    iz = iz > n - 2 ? n - 2 : iz;
    double fx = gx - (double)ix, fy = gy - (double)iy, fz = gz - (double)iz;
    const size_t sy = (size_t)n, sz = (size_t)n * n;
    const float* b = G + ((size_t)iz * sz + (size_t)iy * sy + (size_t)ix);
    double v000 = b[0], v100 = b[1], v010 = b[sy], v110 = b[sy + 1];
    double v001 = b[sz], v101 = b[sz + 1], v011 = b[sz + sy], v111 = b[sz + sy + 1];
    double c00 = v000 + (v100 - v000) * fx;
    double c10 = v010 + (v110 - v010) * fx;
    double c01 = v001 + (v101 - v001) * fx;
    double c11 = v011 + (v111 - v011) * fx;
    double c0 = c00 + (c10 - c00) * fy;
    double c1 = c01 + (c11 - c01) * fy;
    return c0 + (c1 - c0) * fz;
}

__device__ __forceinline__ double turb_factor(const TraceParams& P, double r, double u, double phi, double t) {
    double s = sqrt(fmax(1.0 - u * u, 1e-12));
    double R = r * s, sp, cp;
    bhm::sincos(phi, &sp, &cp);
    double px = R * cp, py = R * sp, pz = r * u;
    if (P.turb_nt <= 1) return grid_trilinear(P, P.turb, px, py, pz);
    // time slices: linear in coordinate time, clamped to the first/last slice
    double st = (t - P.turb_t0) * P.turb_inv_dt;
    st = fmin(fmax(st, 0.0), (double)(P.turb_nt - 1));
    int k = (int)st;
    k = k > P.turb_nt - 2 ? P.turb_nt - 2 : k;
    double f = st - (double)k;
    const size_t slice = (size_t)P.turb_n * P.turb_n * P.turb_n;
    double A = grid_trilinear(P, P.turb + (size_t)k * slice, px, py, pz);
    double B = grid_trilinear(P, P.turb + (size_t)(k + 1) * slice, px, py, pz);
    return A + (B - A) * f;
}

// blackholes/src/emission.cuh:99-175 (exact thermal-synchrotron coefficients)
template <int RADIO>
__device__ __forceinline__ void flow_step(const TraceParams& P, double r, double u, double phi, double t, double lam,
                                          double Ecam, double h, double* dI, double* dtau) {
    *dI = 0.0;
    *dtau = 0.0;
    const double a = P.a;
    double s2 = fmax(1.0 - u * u, 1e-12), s = sqrt(s2);
    double R = r * s, z = r * u;
    if (r > P.r_flow_max || R < 1e-3) return;
    double zh = z / (P.h_r * R);
    double vert = bhm::exp(-0.5 * zh * zh);
    if (vert < 1e-7) return;
    double lr = bhm::log(r);
    double ne = P.n0 * bhm::exp(-P.n_index * lr) * vert;
    if (P.turb_on) {
        if constexpr (RADIO == RADIO_SHEARED) ne *= (double)sheared_factor(P, r, R, z, phi, t);
        else ne *= turb_factor(P, r, u, phi, t);
    }
    ne *= 1.0 - bhm::exp(-(r - P.rh) * (r - P.rh) / 0.25);
    if (ne < 1e-30) return;
    double Th = P.theta0 * bhm::exp(-P.theta_index * lr);
    double B = sqrt(8.0 * PI * ne * C_MP * C_C * C_C / (6.0 * r * P.beta_p));
    double r2 = r * r, a2 = P.a2;
    double Sigma = r2 + a2 * u * u, Delta = r2 - 2.0 * r + a2;
    double A = (r2 + a2) * (r2 + a2) - a2 * Delta * s2;
    double g_tt = -(1.0 - 2.0 * r / Sigma), g_tph = -2.0 * a * r * s2 / Sigma, g_phph = A * s2 / Sigma;
    double omz = 2.0 * a * r / A;
    double Rk = fmax(R, P.r_isco);
    double Om = P.flow_kep * (1.0 / (Rk * sqrt(Rk) + a));
    double vmax = sqrt(Sigma * Delta / A) / sqrt(g_phph) * 0.999;
    Om = omz + fmax(-vmax, fmin(vmax, Om - omz));
    double norm = -(g_tt + 2.0 * g_tph * Om + g_phph * Om * Om);
    if (norm <= 0.0) return;
    double ut = 1.0 / sqrt(norm);
    double kdotu = ut * (1.0 - Om * lam);
    double nu_f = P.nu_obs * Ecam * kdotu;
    double g = P.nu_obs / nu_f;
    double u_ph = ut * (g_tph + g_phph * Om);
    double kb = lam - kdotu * u_ph;
    double bb = g_phph + u_ph * u_ph;
    double cosb = kb / (kdotu * sqrt(bb));
    double sinb = sqrt(fmax(1.0 - cosb * cosb, 1e-6));
    double nuc = C_QE * B / (2.0 * PI * C_ME * C_C);
    double nus = 2.0 / 9.0 * nuc * Th * Th * sinb;
    double X = nu_f / nus;
    double x3 = bhm::cbrt(X);
    double k2 = bhm::exp((double)table_lerp(P.k2, P.k2_n, P.k2_l0, P.k2_dl, (float)bhm::log(Th)));
    double brk = sqrt(X) + 1.887748625363387 * sqrt(x3);
    double j = ne * 1.4142135623730951 * PI * C_QE * C_QE * nus / (3.0 * k2 * C_C) * brk * brk * bhm::exp(-x3);
    double Te = Th * C_ME * C_C * C_C / C_KB;
    double hnu = C_H * nu_f / (C_KB * Te);
    double Bnu = 2.0 * C_H * nu_f * nu_f * nu_f / (C_C * C_C) / bhm::expm1(hnu);
    double alpha = j / Bnu;
    double dl = kdotu * Sigma * h * P.rg_cm;
    double tau = alpha * dl;
    *dtau = tau;
    *dI = g * g * g * Bnu * (-bhm::expm1(-tau));
}

// blackholes/src/kerr.hpp:63-92 (+ trace.cuh:64 direction): initial state and constants
__device__ __forceinline__ void camera_photon(const TraceParams& P, double x, double yv, RayConst& c, State& s) {
    double ox = P.cx + x, oy = P.cy + yv;
    double vx = P.F[0] + P.R[0] * ox + P.U[0] * oy;
    double vy = P.F[1] + P.R[1] * ox + P.U[1] * oy;
    double vz = P.F[2] + P.R[2] * ox + P.U[2] * oy;
    double inv = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
    double kx = -(vx * inv), ky = -(vy * inv), kz = -(vz * inv);
    double k0 = 1.0;
    double b2 = P.beta[0] * P.beta[0] + P.beta[1] * P.beta[1] + P.beta[2] * P.beta[2];
    if (b2 > 0.0) {
        double b = sqrt(b2), g = 1.0 / sqrt(1.0 - b2);
        double ib = 1.0 / b;
        double bx = P.beta[0] * ib, by = P.beta[1] * ib, bz = P.beta[2] * ib;
        double kpar = kx * bx + ky * by + kz * bz;
        double k0z = g * (k0 + b * kpar);
        double f = (g - 1.0) * kpar + g * b * k0;
        kx = kx + bx * f;
        ky = ky + by * f;
        kz = kz + bz * f;
        k0 = k0z;
    }
    const double kr = P.kp[0], ks = P.kp[1], kc = P.kp[2], omega = P.kp[6], e_nu = P.kp[7], e_psi = P.kp[8];
    double L = e_psi * kz;
    double E = e_nu * k0 + omega * L;
    double pth = P.kp[9] * ky;
    double Q = pth * pth + kc * kc * (L * L / (ks * ks) - P.a2 * E * E);
    c.lam = L / E;
    c.eta = Q / (E * E);
    c.E = E;
    c.K = c.eta + (c.lam - P.a) * (c.lam - P.a);
    c.lam2 = c.lam * c.lam;
    c.alam = P.a * c.lam;
    double dr = P.kp[10] * kx / E;
    double dth = pth / E;
    s.y0 = kr;
    s.y1 = -dr;
    s.y2 = kc;
    s.y3 = ks * dth;
    s.y4 = P.ph0;
    s.y5 = P.t0;
}

// blackholes/src/kerr.hpp:175-178
__device__ __forceinline__ double initial_step(const State& s) {
    return 0.02 * fmin(s.y0 / (fabs(s.y1) + 1e-12), 1.0 / (fabs(s.y3) + 1e-3));
}

// Per-ray integration state of trace_ray (trace.cuh:62-80), kept in registers between steps.
struct RayState {
    RayConst c;
    State s;
    double h, acc0, acc1, acc2, acc3, trans, psi, ppx, ppy, ppz;
    int ncross, accepted, flow_steps, it, fate;
    bool done;
};

// blackholes/src/trace.cuh:62-80: initial state of the ray through image-plane offset (x, yv)
template <int RADIO>
__device__ __forceinline__ void ray_init(const TraceParams& P, double x, double yv, RayState& R) {
    camera_photon(P, x, yv, R.c, R.s);
    {
        Rhs k = rhs(P, R.c, R.s.y0, R.s.y2);
        R.s.k1 = k.d1; R.s.k3 = k.d3; R.s.k4 = k.d4; R.s.k5 = k.d5;
    }
    R.h = initial_step(R.s);
    R.acc0 = R.acc1 = R.acc2 = R.acc3 = 0.0;
    R.trans = 1.0;
    R.psi = 0.0;
    R.ncross = R.accepted = R.flow_steps = R.it = R.fate = 0;
    R.ppx = R.ppy = R.ppz = 0.0;
    if (RADIO) {
        double sp = sqrt(fmax(1.0 - R.s.y2 * R.s.y2, 0.0)), sv, cv;
        bhm::sincos(R.s.y4, &sv, &cv);
        R.ppx = sp * cv;
        R.ppy = sp * sv;
        R.ppz = R.s.y2;
    }
    R.done = P.max_steps <= 0;
}

// One iteration of trace.cuh:81-163 (one DP54 attempt); sets R.done when the ray stops.
template <int RADIO>
__device__ __forceinline__ void ray_step(const TraceParams& P, RayState& R) {
    const bool flow = RADIO && P.flow_on;
    State n;
    double hc = R.h;
    if (flow && R.s.y0 < P.r_flow_max) {
        double sn = sqrt(fmax(1.0 - R.s.y2 * R.s.y2, 0.0));
        double cap = P.step_frac * fmin(R.s.y0 / (fabs(R.s.y1) + 1e-30), fmax(sn, 0.05) / (fabs(R.s.y3) + 1e-30));
        hc = fmin(hc, cap);
    }
    double err = dp54_step(P, R.c, R.s, hc, n);
    ++R.it;
    if (!(err <= 1.0)) {
        R.h = hc * fmax(0.2, 0.9 * bhm::pow_m02(fmax(err, 1e-30)));
        if (!(err == err)) R.h = hc * 0.1;
        if (R.it >= P.max_steps) R.done = true;
        return;
    }
    ++R.accepted;
    if (RADIO) {
        double sn = sqrt(fmax(1.0 - n.y2 * n.y2, 0.0)), sv, cv;
        bhm::sincos(n.y4, &sv, &cv);
        double pnx = sn * cv, pny = sn * sv, pnz = n.y2;
        double cx = R.ppy * pnz - R.ppz * pny, cy = R.ppz * pnx - R.ppx * pnz, cz = R.ppx * pny - R.ppy * pnx;
        double cl = sqrt(cx * cx + cy * cy + cz * cz);
        double dt = R.ppx * pnx + R.ppy * pny + R.ppz * pnz;
        R.psi += bhm::atan2(cl, dt);
        R.ppx = pnx;
        R.ppy = pny;
        R.ppz = pnz;
    }
    if ((R.s.y2 > 0.0) != (n.y2 > 0.0)) ++R.ncross;
    if (flow && n.y0 < P.r_flow_max) {
        double dI, dtau;
        ++R.flow_steps;
        flow_step<RADIO>(P, n.y0, n.y2, n.y4, n.y5, R.c.lam, R.c.E, hc, &dI, &dtau);
        double cI = R.trans * dI;
        int ord = (int)(R.psi * (1.0 / PI));
        R.acc0 += cI;
        // value selects, not a branch chain: the chain becomes a dynamically indexed stack array
        R.acc1 += ord == 1 ? cI : 0.0;
        R.acc2 += ord == 2 ? cI : 0.0;
        R.acc3 += ord >= 3 ? cI : 0.0;
        R.trans *= bhm::exp(-dtau);
        if (R.trans < 1e-5) {
            R.fate = -1;
            R.done = true;
            return;
        }
    }
    R.s = n;
    project_radial(P, R.c, R.s);
    R.h = hc * fmin(5.0, fmax(0.2, 0.9 * bhm::pow_m02(fmax(err, 1e-30))));
    if (R.s.y0 < P.r_stop) {
        R.fate = -1;
        R.done = true;
        return;
    }
    if (R.s.y0 > P.r_escape && R.s.y1 > 0.0) {
        R.fate = 1;   // escape angles are taken from the final state in ray_result
        R.done = true;
        return;
    }
    if (R.it >= P.max_steps) R.done = true;
}

// blackholes/src/trace.cuh:25-33: galactic texture coordinates of the escape direction (th, ph),
// the direction rotated by sky_rot (black-hole frame -> galactic) and taken to (longitude, latitude)
__device__ __forceinline__ void sky_uv(const TraceParams& P, double th, double ph, float* u, float* v) {
    double st, ct, sp, cp;
    bhm::sincos(th, &st, &ct);
    bhm::sincos(ph, &sp, &cp);
    const double dx = st * cp, dy = st * sp, dz = ct;
    const double gx = P.sky_rot[0] * dx + P.sky_rot[1] * dy + P.sky_rot[2] * dz;
    const double gy = P.sky_rot[3] * dx + P.sky_rot[4] * dy + P.sky_rot[5] * dz;
    const double gz = P.sky_rot[6] * dx + P.sky_rot[7] * dy + P.sky_rot[8] * dz;
    const double l = bhm::atan2(gy, gx), b = bhm::asin(fmax(-1.0, fmin(1.0, gz)));
    float uu = (float)(0.5 - l * (0.5 / PI));
    uu -= floorf(uu);
    *u = uu;
    *v = (float)(0.5 - b / PI);
}

// Radiance of the procedural sky at (u, v): sky_nu x sky_nv squares, bright sky_scale, dark 0.1 of it
// (the reference samples its star-map texture here, trace.cuh:35-45)
__device__ __forceinline__ float sky_checker(const TraceParams& P, float u, float v) {
    const int iu = min(max((int)(u * (float)P.sky_nu), 0), P.sky_nu - 1);
    const int iv = min(max((int)(v * (float)P.sky_nv), 0), P.sky_nv - 1);
    return ((iu + iv) & 1) ? P.sky_scale : 0.1f * P.sky_scale;
}

// Visible mode, escaped ray (trace.cuh:154-160): sky coordinates and, with the sky on, its radiance
// in the first three channels (the reference's X, Y, Z), weighted by the transmission
__device__ __forceinline__ void sky_result(const TraceParams& P, double trans, RayResult& out) {
    out.sky_u = 0.f;
    out.sky_v = 0.f;
    if (P.mode != MODE_VISIBLE || out.fate != 1) return;
    sky_uv(P, out.th_inf, out.ph_inf, &out.sky_u, &out.sky_v);
    if (P.sky_on) {
        const double s = (double)sky_checker(P, out.sky_u, out.sky_v);
        out.I[0] += trans * s;
        out.I[1] += trans * s;
        out.I[2] += trans * s;
    }
}

__device__ __forceinline__ void ray_result(const TraceParams& P, const RayState& R, RayResult& out) {
    out.y[0] = R.s.y0; out.y[1] = R.s.y1; out.y[2] = R.s.y2; out.y[3] = R.s.y3; out.y[4] = R.s.y4;
    out.y[5] = R.s.y5;
    out.th_inf = 0.0;
    out.ph_inf = 0.0;
    if (R.fate == 1) {
        // blackholes/src/trace.cuh:49-55 (k1[4] is the phi derivative of the final state)
        double sq = sqrt(fmax(1.0 - R.s.y2 * R.s.y2, 1e-30));
        double tq = bhm::acos(fmax(-1.0, fmin(1.0, R.s.y2)));
        double dthdr = (-R.s.y3 / sq) / R.s.y1, dphdr = R.s.k4 / R.s.y1;
        out.th_inf = tq + R.s.y0 * dthdr;
        out.ph_inf = R.s.y4 + R.s.y0 * dphdr;
    }
    out.I[0] = R.acc0; out.I[1] = R.acc1; out.I[2] = R.acc2; out.I[3] = R.acc3;
    out.trans = R.trans;
    out.psi = R.psi;
    out.fate = R.fate;
    out.ncross = R.ncross;
    out.order = (int)(R.psi * (1.0 / PI));
    out.steps = R.it;
    out.accepted = R.accepted;
    out.flow_steps = R.flow_steps;
    sky_result(P, R.trans, out);
}

// blackholes/src/trace.cuh:62-165 (disk branch out of scope, procedural sky), FP64 intensity.
template <int RADIO>
__device__ __forceinline__ void trace_ray_fp64(const TraceParams& P, double x, double yv, RayResult& out) {
    RayState R;
    ray_init<RADIO>(P, x, yv, R);
    while (!R.done) ray_step<RADIO>(P, R);
    ray_result(P, R, out);
}

// blackholes/src/kerr.hpp:184-203 for one photon (lam, eta, E, y[6])
__device__ __forceinline__ int vacuum_fate(const TraceParams& P, const double* ph, double tol, double r_escape,
                                           int max_steps) {
    RayConst c;
    c.lam = ph[0];
    c.eta = ph[1];
    c.E = ph[2];
    c.K = c.eta + (c.lam - P.a) * (c.lam - P.a);
    c.lam2 = c.lam * c.lam;
    c.alam = P.a * c.lam;
    State s, n;
    s.y0 = ph[3]; s.y1 = ph[4]; s.y2 = ph[5]; s.y3 = ph[6]; s.y4 = ph[7]; s.y5 = ph[8];
    Rhs k = rhs(P, c, s.y0, s.y2);
    s.k1 = k.d1; s.k3 = k.d3; s.k4 = k.d4; s.k5 = k.d5;
    double h = initial_step(s);
    const double rcap = P.rh * 1.0001 + 1e-3;
    for (int it = 0; it < max_steps; ++it) {
        double err = dp54_step(P, c, s, h, n);
        if (err > 1.0) {
            h *= fmax(0.2, 0.9 * bhm::pow_m02(err));
            continue;
        }
        s = n;
        project_radial(P, c, s);
        h *= fmin(5.0, fmax(0.2, 0.9 * bhm::pow_m02(fmax(err, 1e-30))));
        if (s.y0 < rcap) return -1;
        if (s.y0 > r_escape && s.y1 > 0.0) return 1;
    }
    return 0;
}

}  // namespace bhd
