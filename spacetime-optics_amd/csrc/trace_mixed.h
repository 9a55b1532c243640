/* trace_mixed.h - mixed FP32/FP64 per-ray tracer for gfx950 (device code, included by bhrt.hip).

   Policy (chosen from CPU sweeps against the FP64 oracle on the 1024^2 230 GHz frame and from
   timings on MI355X):
   - Tolerance regions: the scene tolerance inside r_tol (radio: the emission region, the radius
     holding all but 1e-6 of the flow's emission with a negligible optical depth outside; vacuum:
     r_near, near the hole) - there the reference's steps are tolerance-limited and the
     radiative-transfer quadrature depends on that step sequence; tol_mid from r_tol to r_far (the
     flow's edge in radio mode); tol_far beyond (radio: scaled with the pixel's size, 1e-4 on
     0.04 M pixels, and by sin(theta_cam) / sin(15 deg) for a camera within 15 degrees of the spin
     axis, where the norm's bound on cos(theta) is a looser bound on theta; vacuum: 1e-6). Outside
     r_tol the step-size controller also scales the next step
     by r_start / r_end of the accepted one (the 1/r scaling of the Mino-time step far from the
     hole): the reference's error-only controller lags behind it and rejected every other step on
     the way out (CPU sweep: 114 -> 95 attempts per ray, rejections 18 -> 2.6, errors unchanged).
     Inside r_tol the controller is the reference's.
   - Emission region: emission, absorption and the transfer step cap are evaluated below r_cut =
     r_tol; the rest of the flow is traced as vacuum with tol_mid = tol_far / 10 (or, with
     full_flow, the whole flow as the FP64 tracer with tol_mid = 1e-6).
   - Stage precision: FP64 stages evaluated with fused multiply-adds and Newton-refined FP64
     reciprocals; optionally (r >= r_f32) DP5(4) stages in FP32 with hardware reciprocals, the
     state accumulated in FP64, the error norm, step controller and radial-constraint projection
     in FP64. The coordinate time is integrated only where something reads it (time-dependent
     turbulence, a t bound, ray records). The projection r' = +-sqrt(R(r)) after every accepted
     step keeps the radial motion on the photon's constant-of-motion surface, so FP32 stage errors
     only shift where along the trajectory a step lands. The FP32 stages are off by default:
     without packed FP32 the arithmetic runs at the FP64 rate, and lanes of a wave crossing r_f32
     at different steps run both stage codes.
   - Emission and absorption in FP32, in log space (the cgs coefficients span 1e-45..1e-10), exact
     thermal-synchrotron formula or the MLP surrogate table; intensity and transmission accumulate
     in FP64. The swept angle (photon-ring order) is tracked in FP32. In the persistent engine each
     ray keeps its last turbulence cell in registers, so consecutive samples in one cell load
     nothing; the surrogate's table loads are in flight while the exact formula runs.
   None of this is bit-compatible with the FP64 tracer; its error against the FP64 oracle is what
   the tests bound. */
#pragma once
#include "mlp.h"
#include "trace_fp64.h"

namespace bhm32 {

using bhd::RayConst;
using bhd::RayResult;
using bhd::State;
using bhd::TraceParams;

// Multiply-adds fuse within one source expression only (not across statements): which operations
// fuse is then fixed by the source, not by the inlining context, so a ray gives the same bits in
// every kernel that runs it (frame kernel, persistent kernel).
#pragma clang fp contract(on)

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// hardware FP32 square root (1.5 ulp): the correctly rounded sqrtf is a 15-instruction sequence
__device__ __forceinline__ float sqrtf_(float x) { return __builtin_amdgcn_sqrtf(x); }
// v_rcp_f64 is good to 2^-24.4 relative on MI355X; one Newton step brings it to 2.2e-15 (measured
// over 2^20 values spanning 1e-13..1e13)
__device__ __forceinline__ double rcpd(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

template <typename R>
__device__ __forceinline__ R rcp(R x);
template <>
__device__ __forceinline__ float rcp<float>(float x) { return rcpf(x); }
template <>
__device__ __forceinline__ double rcp<double>(double x) { return rcpd(x); }

// natural log / exp on the hardware base-2 units
__device__ __forceinline__ float lnf(float x) { return __builtin_amdgcn_logf(x) * 0.69314718055994531f; }
__device__ __forceinline__ float expf_(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// x^-0.2 in FP64 (the step-size controller's factor) for x clamped to [1e-30, 1e30]: an FP32 seed from
// the hardware log2 / exp2 (relative error ~2e-7), then one Newton step on x y^5 = 1,
// y <- y + y (1 - x y^5) / 5, which squares the relative error (times 3): ~1e-13.
// A NaN x gives the factor of 1e-30 (the callers handle a NaN error themselves).
__device__ __forceinline__ double pow_m02_d(double x) {
    const double xc = fmin(fmax(x, 1e-30), 1e30);
    const double y = (double)__builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf((float)xc));
    const double y2 = y * y, y5 = y2 * y2 * y;
    return fma(y * 0.2, fma(-xc, y5, 1.0), y);
}

// FP64 square root of x >= 0 from the hardware reciprocal square root and two Goldschmidt
// refinements (a few ulp; the correctly rounded sqrt adds scaling and fix-up for ranges the tracer
// never produces: x is a radial potential of order r^4, far from the denormals)
__device__ __forceinline__ double sqrt_d(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    return x > 0.0 ? fma(d, h, g) : 0.0;
}

// blackholes/src/kerr.hpp:166-173 with the square root above
__device__ __forceinline__ void project_radial_m(const TraceParams& P, const RayConst& c, State& s) {
    const double r = s.y0;
    const double Pp = r * r + P.a2 - c.alam, D = r * r - 2.0 * r + P.a2;
    const double R = Pp * Pp - D * c.K;
    if (R > 0.25 * r * r * r * r) {
        const double v = sqrt_d(R);
        s.y1 = s.y1 < 0.0 ? -v : v;
    }
}

// asin on [0, 1] in FP32: t + t^3 P(t^2) on [0, 1/2] (P of degree 4, least squares on Chebyshev
// nodes for the absolute error: 1.6e-9, 4.7e-8 with FP32 evaluation), and
// pi/2 - 2 asin(sqrt((1 - x) / 2)) above 1/2
__device__ __forceinline__ float asin_pos(float x) {
    const bool big = x > 0.5f;
    const float t = big ? sqrtf_(0.5f * (1.f - x)) : x, z = t * t;
    float p = 0.0437449403f;
    p = p * z + 0.0231502242f;
    p = p * z + 0.0457071587f;
    p = p * z + 0.0749306679f;
    p = p * z + 0.166668221f;
    const float a = t + t * z * p;
    return big ? 1.5707963267948966f - 2.f * a : a;
}

template <typename R>
struct Rhs {
    R d1, d3, d4, d5;
};

template <typename R>
struct RConst {
    R a, a2, lam, eta, K, lam2, alam, s2min;
};

// blackholes/src/kerr.hpp:103-116 with reciprocals in place of the three divisions; without TIME
// the coordinate-time derivative d5 is left at zero
template <typename R, bool TIME = true>
__device__ __forceinline__ Rhs<R> rhs(const RConst<R>& c, R r, R u) {
    R r2 = r * r;
    R Delta = r2 - R(2) * r + c.a2;
    R Pp = r2 + c.a2 - c.alam;
    R s2 = fmax(R(1) - u * u, c.s2min);
    R iD = rcp<R>(Delta), is2 = rcp<R>(s2);
    Rhs<R> o;
    o.d1 = R(2) * r * Pp - (r - R(1)) * c.K;
    o.d3 = u * (c.a2 * (R(1) - R(2) * u * u) - c.eta - c.lam2);
    o.d4 = -(c.a * Pp * iD + c.lam * is2 - c.a);
    o.d5 = TIME ? -((r2 + c.a2) * Pp * iD + c.a * (c.lam - c.a * s2)) : R(0);
    return o;
}

/* One Dormand-Prince 5(4) step with stage arithmetic in R (float: far field, double: near field),
   the 5th-order state accumulated in FP64 and the scaled error norm in FP64. Without TIME the
   coordinate time t and its derivative are carried over unchanged: t is in no component of the
   reference's error norm, so the step sequence and every other component are the same bits. */
template <typename R, bool TIME = true>
__device__ __forceinline__ double dp54(const RConst<R>& c, const State& s, double h, double inv_tol, double inv_tol_t,
                                       State& o) {
    using namespace bhd;
    const R hr = (R)h;
    const R y0 = (R)s.y0, y1 = (R)s.y1, y2 = (R)s.y2, y3 = (R)s.y3;
    R q0 = y1, q1 = (R)s.k1, q2 = y3, q3 = (R)s.k3;
    R S3_0 = R(a31) * q0, S3_1 = R(a31) * q1, S3_2 = R(a31) * q2, S3_3 = R(a31) * q3;
    R S4_0 = R(a41) * q0, S4_1 = R(a41) * q1, S4_2 = R(a41) * q2, S4_3 = R(a41) * q3;
    R S5_0 = R(a51) * q0, S5_1 = R(a51) * q1, S5_2 = R(a51) * q2, S5_3 = R(a51) * q3;
    R S6_0 = R(a61) * q0, S6_1 = R(a61) * q1, S6_2 = R(a61) * q2, S6_3 = R(a61) * q3;
    R O0 = R(b1) * q0, O1 = R(b1) * q1, O2 = R(b1) * q2, O3 = R(b1) * q3, O4 = R(b1) * (R)s.k4,
      O5 = TIME ? R(b1) * (R)s.k5 : R(0);
    R E0 = R(e1) * q0, E1 = R(e1) * q1, E2 = R(e1) * q2, E3 = R(e1) * q3, E4 = R(e1) * (R)s.k4,
      E5 = TIME ? R(e1) * (R)s.k5 : R(0);
    const R ha = hr * R(a21);
    R t0 = y0 + ha * q0, t1 = y1 + ha * q1, t2 = y2 + ha * q2, t3 = y3 + ha * q3;
    Rhs<R> k = rhs<R, TIME>(c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S3_0 += R(a32) * q0; S3_1 += R(a32) * q1; S3_2 += R(a32) * q2; S3_3 += R(a32) * q3;
    S4_0 += R(a42) * q0; S4_1 += R(a42) * q1; S4_2 += R(a42) * q2; S4_3 += R(a42) * q3;
    S5_0 += R(a52) * q0; S5_1 += R(a52) * q1; S5_2 += R(a52) * q2; S5_3 += R(a52) * q3;
    S6_0 += R(a62) * q0; S6_1 += R(a62) * q1; S6_2 += R(a62) * q2; S6_3 += R(a62) * q3;
    t0 = y0 + hr * S3_0; t1 = y1 + hr * S3_1; t2 = y2 + hr * S3_2; t3 = y3 + hr * S3_3;
    k = rhs<R, TIME>(c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S4_0 += R(a43) * q0; S4_1 += R(a43) * q1; S4_2 += R(a43) * q2; S4_3 += R(a43) * q3;
    S5_0 += R(a53) * q0; S5_1 += R(a53) * q1; S5_2 += R(a53) * q2; S5_3 += R(a53) * q3;
    S6_0 += R(a63) * q0; S6_1 += R(a63) * q1; S6_2 += R(a63) * q2; S6_3 += R(a63) * q3;
    O0 += R(b3) * q0; O1 += R(b3) * q1; O2 += R(b3) * q2; O3 += R(b3) * q3; O4 += R(b3) * k.d4;
    E0 += R(e3) * q0; E1 += R(e3) * q1; E2 += R(e3) * q2; E3 += R(e3) * q3; E4 += R(e3) * k.d4;
    if (TIME) { O5 += R(b3) * k.d5; E5 += R(e3) * k.d5; }
    t0 = y0 + hr * S4_0; t1 = y1 + hr * S4_1; t2 = y2 + hr * S4_2; t3 = y3 + hr * S4_3;
    k = rhs<R, TIME>(c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S5_0 += R(a54) * q0; S5_1 += R(a54) * q1; S5_2 += R(a54) * q2; S5_3 += R(a54) * q3;
    S6_0 += R(a64) * q0; S6_1 += R(a64) * q1; S6_2 += R(a64) * q2; S6_3 += R(a64) * q3;
    O0 += R(b4) * q0; O1 += R(b4) * q1; O2 += R(b4) * q2; O3 += R(b4) * q3; O4 += R(b4) * k.d4;
    E0 += R(e4) * q0; E1 += R(e4) * q1; E2 += R(e4) * q2; E3 += R(e4) * q3; E4 += R(e4) * k.d4;
    if (TIME) { O5 += R(b4) * k.d5; E5 += R(e4) * k.d5; }
    t0 = y0 + hr * S5_0; t1 = y1 + hr * S5_1; t2 = y2 + hr * S5_2; t3 = y3 + hr * S5_3;
    k = rhs<R, TIME>(c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    S6_0 += R(a65) * q0; S6_1 += R(a65) * q1; S6_2 += R(a65) * q2; S6_3 += R(a65) * q3;
    O0 += R(b5) * q0; O1 += R(b5) * q1; O2 += R(b5) * q2; O3 += R(b5) * q3; O4 += R(b5) * k.d4;
    E0 += R(e5) * q0; E1 += R(e5) * q1; E2 += R(e5) * q2; E3 += R(e5) * q3; E4 += R(e5) * k.d4;
    if (TIME) { O5 += R(b5) * k.d5; E5 += R(e5) * k.d5; }
    t0 = y0 + hr * S6_0; t1 = y1 + hr * S6_1; t2 = y2 + hr * S6_2; t3 = y3 + hr * S6_3;
    k = rhs<R, TIME>(c, t0, t2);
    q0 = t1; q1 = k.d1; q2 = t3; q3 = k.d3;
    O0 += R(b6) * q0; O1 += R(b6) * q1; O2 += R(b6) * q2; O3 += R(b6) * q3; O4 += R(b6) * k.d4;
    E0 += R(e6) * q0; E1 += R(e6) * q1; E2 += R(e6) * q2; E3 += R(e6) * q3; E4 += R(e6) * k.d4;
    if (TIME) { O5 += R(b6) * k.d5; E5 += R(e6) * k.d5; }
    o.y0 = s.y0 + (double)(hr * O0);
    o.y1 = s.y1 + (double)(hr * O1);
    o.y2 = s.y2 + (double)(hr * O2);
    o.y3 = s.y3 + (double)(hr * O3);
    o.y4 = s.y4 + (double)(hr * O4);
    o.y5 = TIME ? s.y5 + (double)(hr * O5) : s.y5;
    k = rhs<R, TIME>(c, (R)o.y0, (R)o.y2);
    o.k1 = (double)k.d1;
    o.k3 = (double)k.d3;
    o.k4 = (double)k.d4;
    o.k5 = TIME ? (double)k.d5 : s.k5;
    E0 += R(e7) * (R)o.y1; E1 += R(e7) * k.d1; E2 += R(e7) * (R)o.y3; E3 += R(e7) * k.d3; E4 += R(e7) * k.d4;
    if (TIME) E5 += R(e7) * k.d5;
    // scaled max-norm over r, r', u, u', phi (kerr.hpp:145-155) in FP64. Each of its scales is tol
    // times a factor a_i (a = 1 for u, 4 for phi), so the norm is max_i(|h E_i| / a_i) / tol: three
    // Newton-refined FP64 reciprocals for the state-dependent factors, and 1 / tol precomputed
    const double a0 = 1e-6 + fmax(fabs(s.y0), fabs(o.y0));
    const double a1 = 1e-6 + fmax(fmax(fabs(s.y1), fabs(o.y1)), s.y0 * s.y0);
    const double a3 = fmax(1.0, fmax(fabs(s.y3), fabs(o.y3)));
    double m = fabs(h * (double)E0) * rcpd(a0);
    m = fmax(m, fabs(h * (double)E1) * rcpd(a1));
    m = fmax(m, fabs(h * (double)E2));
    m = fmax(m, fabs(h * (double)E3) * rcpd(a3));
    m = fmax(m, fabs(h * (double)E4) * 0.25);
    m *= inv_tol;
    // coordinate time (not in the reference's norm): an absolute bound per step where the turbulence
    // depends on the emission time (inv_tol_t = 1 / bound, 0 when it does not)
    if (TIME) m = fmax(m, fabs(h * (double)E5) * inv_tol_t);
    return m;
}

template <typename R>
__device__ __forceinline__ RConst<R> make_const(const TraceParams& P, const RayConst& c) {
    RConst<R> k;
    k.a = (R)P.a;
    k.a2 = (R)P.a2;
    k.lam = (R)c.lam;
    k.eta = (R)c.eta;
    k.K = (R)c.K;
    k.lam2 = (R)c.lam2;
    k.alam = (R)c.alam;
    k.s2min = sizeof(R) == 4 ? (R)1e-30 : (R)1e-300;
    return k;
}

// ---------------------------------------------------------------------------------------------
// Emission (blackholes/src/emission.cuh:99-175) in FP32, log space.
struct EmisParams {
    // log-space constants, filled on the host
    float ln_n0, n_index, ln_theta0, theta_index, h_r, inv_h_r, r_flow_max, rh, r_isco, flow_kep, a, a2;
    float ln_bconst;   // ln(8 pi m_p c^2 / (6 beta)) : B^2 = exp(ln_bconst) n_e / r
    float ln_nuc_per_B;   // ln(e / (2 pi m_e c))
    float ln_jconst;   // ln(sqrt(2) pi e^2 / (3 c))
    float ln_bnu_const;   // ln(2 h / c^2)
    float ln_h_over_mec2; // ln(h / (m_e c^2))  (x = h nu / (k T_e) = h nu / (m_e c^2 Theta))
    float ln_nu_obs;
    float ln_rg;
    float k2_inv_dl;   // 1 / (K2 table spacing)
};

// the K2 table lookup of bhd::table_lerp with the spacing's reciprocal in place of the division
__device__ __forceinline__ float table_lerp_f32(const float* __restrict__ t, int n, float x0, float inv_dx, float x) {
    float f = (x - x0) * inv_dx;
    f = fminf(fmaxf(f, 0.f), (float)(n - 2) + 0.9999f);
    const int i = (int)f;
    f -= (float)i;
    const float t0 = t[i], t1 = t[i + 1];
    return t0 + (t1 - t0) * f;
}

// Thermal synchrotron: ln j, ln alpha and ln B_nu from ln n_e, ln Theta, ln nu_f, ln nu_c,
// ln sin(theta_B) (Leung et al. 2011 with the K2 table, Kirchhoff alpha = j / B_nu). With
// y = X^(1/3): sqrt(X) + 2^(11/12) X^(1/6) = X^(1/6) (y + 2^(11/12)), so ln j needs one exponential
// and one logarithm; X is capped at e^240 (where j = exp(-y) is already 0 in FP32) so y stays finite.
__device__ __forceinline__ void synchrotron_exact(const EmisParams& E, float ln_ne, float ln_th, float lnk2,
                                                  float ln_nu, float ln_nuc, float ln_sinb, float* ln_j,
                                                  float* ln_a, float* ln_bnu) {
    const float ln_nus = -1.5040774f + ln_nuc + 2.f * ln_th + ln_sinb;   // ln(2/9) + ...
    const float lnX = fminf(ln_nu - ln_nus, 240.f);
    const float y = expf_(lnX * (1.f / 3.f));
    *ln_j = ln_ne + E.ln_jconst + ln_nus - lnk2 + lnX * (1.f / 3.f) + 2.f * lnf(y + 1.887748625363387f) - y;
    // B_nu with x = h nu / (m_e c^2 Theta): ln(2 h nu^3 / c^2) - ln(expm1(x)); ln x is known
    const float ln_x = E.ln_h_over_mec2 + ln_nu - ln_th;
    // ln(expm1(x)) = ln x + x/2 + ...; for ln x < -15 the x/2 term is below half an FP32 ulp of ln x,
    // so the sum is ln x itself and x need not be formed (230 GHz in this flow: x ~ 1e-9)
    float ln_em1 = ln_x;
    if (ln_x >= -15.f) {
        const float x = expf_(ln_x);
        ln_em1 = x < 1e-3f ? ln_x + 0.5f * x : (x > 30.f ? x : lnf(expf_(x) - 1.f));
    }
    *ln_bnu = E.ln_bnu_const + 3.f * ln_nu - ln_em1;
    *ln_a = *ln_j - *ln_bnu;
}

// Turbulence density factor in FP32: the grid, cell-centred trilinear weights, edge clamps and
// time-slice interpolation of bhd::turb_factor, with the azimuth reduced to [0, 2 pi) in FP64 before
// the FP32 sine and cosine (an orbiting photon's phi grows by 2 pi per turn). Split in two: the fetch
// computes the cell and issues the loads of its corners (one slice, or the two slices around the
// emission time); the blend interpolates them. Between the two the tracer computes what does not
// depend on the density, so the loads' latency is hidden behind that work. In the persistent engine
// (CACHE) each ray keeps the corners of its last cell (slice k) in registers: consecutive samples of
// a ray usually fall in the same cell (steps of 0.1-0.4 M against ~1 M cells near the hole), and
// then the fetch loads nothing, so a ray running alone in its wave does not wait for memory. The
// values are those in memory, so the result is the same bits either way.
struct Corners {
    float c000, c100, c010, c110, c001, c101, c011, c111;
};
struct TurbCache {
    Corners v;          // the corners of cell `key` (slice k)
    unsigned int key;   // grid offset of the cell's first corner, slice included; ~0u: none
};
struct TurbFetch {
    Corners k1;   // slice k + 1 (time slices only)
    float fx, fy, fz, ft;
    bool two;
};

__device__ __forceinline__ Corners grid_corners_f32(const float* __restrict__ b, unsigned int sy, unsigned int sz) {
    return Corners{b[0], b[1], b[sy], b[sy + 1], b[sz], b[sz + 1], b[sz + sy], b[sz + sy + 1]};
}

__device__ __forceinline__ float grid_blend_f32(const Corners& v, float fx, float fy, float fz) {
    const float c00 = v.c000 + (v.c100 - v.c000) * fx, c10 = v.c010 + (v.c110 - v.c010) * fx;
    const float c01 = v.c001 + (v.c101 - v.c001) * fx, c11 = v.c011 + (v.c111 - v.c011) * fx;
    const float c0 = c00 + (c10 - c00) * fy, c1 = c01 + (c11 - c01) * fy;
    return c0 + (c1 - c0) * fz;
}

template <bool CACHE>
__device__ __forceinline__ void turb_fetch_f32(const TraceParams& P, float R, float r, float u, double phid,
                                               double td, TurbFetch& T, TurbCache& C) {
    const double two_pi = 6.283185307179586;
    const float ph = (float)(phid - two_pi * floor(phid * (1.0 / two_pi)));
    float sp, cp;
    __sincosf(ph, &sp, &cp);
    const float px = R * cp, py = R * sp, pz = r * u;
    const int n = P.turb_n;
    const float half = (float)P.turb_half, inv = (float)P.turb_inv_dx, hi = (float)(n - 1);
    const float gx = fminf(fmaxf((px + half) * inv - 0.5f, 0.f), hi);
    const float gy = fminf(fmaxf((py + half) * inv - 0.5f, 0.f), hi);
    const float gz = fminf(fmaxf((pz + half) * inv - 0.5f, 0.f), hi);
    const int ix = min((int)gx, n - 2), iy = min((int)gy, n - 2), iz = min((int)gz, n - 2);
    T.fx = gx - (float)ix;
    T.fy = gy - (float)iy;
    T.fz = gz - (float)iz;
    // 32-bit offsets: bh_create keeps the grid (n^3 cells x nt slices) below 2^31 cells
    const unsigned int sy = (unsigned int)n, sz = (unsigned int)(n * n);
    const unsigned int cell = (unsigned int)iz * sz + (unsigned int)iy * sy + (unsigned int)ix;
    T.two = P.turb_nt > 1;
    // the slice (pair) and its weight first; the corner loads then write the same fields on both
    // paths, so the structs stay in registers (no store through a selected address)
    unsigned int key = cell, slice = 0;
    T.ft = 0.f;
    if (T.two) {
        const float st = fminf(fmaxf((float)((td - P.turb_t0) * P.turb_inv_dt), 0.f), (float)(P.turb_nt - 1));
        const int k = min((int)st, P.turb_nt - 2);
        T.ft = st - (float)k;
        slice = (unsigned int)(P.turb_n * P.turb_n * P.turb_n);
        key += (unsigned int)k * slice;
    }
    if (!CACHE || key != C.key) {   // a new cell: the loads refill the ray's cached corners in place
        C.v = grid_corners_f32(P.turb + key, sy, sz);
        C.key = key;
    }
    if (T.two) T.k1 = grid_corners_f32(P.turb + key + slice, sy, sz);   // read only when two
}

__device__ __forceinline__ float turb_blend_f32(const TurbFetch& T, const TurbCache& C) {
    const float A = grid_blend_f32(C.v, T.fx, T.fy, T.fz);
    if (!T.two) return A;
    const float B = grid_blend_f32(T.k1, T.fx, T.fy, T.fz);
    return A + (B - A) * T.ft;
}

// The sheared turbulence in FP32 (same slices and filter as bhd::sheared_factor): slice coordinates
// from FP32 logs, the emission time relative to the camera's taken in FP64 before it is rounded.
__device__ __forceinline__ float sheared_factor_f32(const TraceParams& P, float r, float R, float z, double phid,
                                                    double td) {
    const bhd::ShearedView& T = *P.ts;
    const float u = (lnf(R) - T.lr0) * rcpf(T.lr1 - T.lr0);
    const float v = (float)(phid * (0.5 / bhd::PI));
    const float w = 0.5f + 0.5f * (z * rcpf(R)) * rcpf(T.zmax);
    const float cr = (float)T.cam_r;
    const float dt = (float)(td - T.cam_t) + (cr - r + 2.f * lnf(cr * rcpf(r)));   // t - t_ref(r)
    const float s = fminf(fmaxf((dt - T.d0) * rcpf(T.dd), 0.f), (float)(T.nt - 1) - 1e-4f);
    const int k = (int)s;
    const float f = s - (float)k;
    const size_t slice = (size_t)T.nphi * T.nlr * T.nz;
    const float A = bhd::sh_tex(T.data + (size_t)k * slice, T.nphi, T.nlr, T.nz, v, u, w);
    const float B = bhd::sh_tex(T.data + (size_t)(k + 1) * slice, T.nphi, T.nlr, T.nz, v, u, w);
    return A + (B - A) * f;
}

// Geometry + plasma of one flow sample, in two parts. flow_begin takes everything that does not
// depend on the density - the vertical profile, the horizon softening, the fluid four-velocity, the
// Doppler factor, the pitch angle, the path length - and issues the turbulence loads (grid model);
// flow_finish blends them into the density and the cyclotron frequency. Each returns false when the
// sample emits and absorbs nothing; the early-outs are those of emission.cuh:99-175.
struct FlowSample {
    float ln_ne, ln_th, ln_nu, ln_nuc, ln_sinb, ln_g, ln_dl;
    float ln_k2;   // ln K2(1 / Theta_e) from the table: looked up early, its loads overlap the rest
};
struct FlowPending {
    float lr, lnvert, soft, R, z;
    TurbFetch tf;
};

template <int RADIO, bool TCACHE>
__device__ __forceinline__ bool flow_begin(const TraceParams& P, const EmisParams& E, double rd, double ud,
                                           double phid, double td, double lamd, double Ecamd, double hd,
                                           FlowSample& f, FlowPending& q, TurbCache& C) {
    const float r = (float)rd, u = (float)ud, lam = (float)lamd;
    // sin theta and its reciprocal from one reciprocal square root
    const float s2 = fmaxf(1.f - u * u, 1e-12f), is = __builtin_amdgcn_rsqf(s2), s = s2 * is;
    const float R = r * s, z = r * u;
    if (r > E.r_flow_max || R < 1e-3f) return false;
    const float zh = u * is * E.inv_h_r;   // z / (H/R R) = u / (H/R sin theta)
    const float lnvert = -0.5f * zh * zh;
    if (lnvert < -16.118096f) return false;   // vert < 1e-7
    if constexpr (RADIO == bhd::RADIO_GRID)
        if (P.turb_on) turb_fetch_f32<TCACHE>(P, R, r, u, phid, td, q.tf, C);   // loads in flight until flow_finish
    const float lr = lnf(r);
    const float dr = r - E.rh;
    const float soft = 1.f - expf_(-dr * dr * 4.f);
    if (!(soft > 0.f)) return false;
    const float ln_th = E.ln_theta0 - E.theta_index * lr;
    f.ln_k2 = table_lerp_f32(P.k2, P.k2_n, P.k2_l0, E.k2_inv_dl, ln_th);
    // fluid four-velocity (emission.cuh:138-149)
    const float r2 = r * r, a = E.a, a2 = E.a2;
    const float Sigma = r2 + a2 * u * u, Delta = r2 - 2.f * r + a2;
    const float A = (r2 + a2) * (r2 + a2) - a2 * Delta * s2;
    const float iSA = rcpf(Sigma * A), iS = A * iSA, iA = Sigma * iSA;   // 1 / Sigma, 1 / A from one reciprocal
    const float g_tt = -(1.f - 2.f * r * iS), g_tph = -2.f * a * r * s2 * iS, g_phph = A * s2 * iS;
    const float omz = 2.f * a * r * iA;
    const float Rk = fmaxf(R, E.r_isco);
    float Om = E.flow_kep * rcpf(Rk * sqrtf_(Rk) + a);
    // sqrt(Sigma Delta / A) / sqrt(g_phph) = Sigma sqrt(Delta) / (A sin theta)
    const float vmax = Sigma * iA * is * sqrtf_(Delta) * 0.999f;
    Om = omz + fmaxf(-vmax, fminf(vmax, Om - omz));
    const float norm = -(g_tt + 2.f * g_tph * Om + g_phph * Om * Om);
    if (!(norm > 0.f)) return false;
    const float ut = __builtin_amdgcn_rsqf(norm);
    const float kdotu = ut * (1.f - Om * lam);
    if (!(kdotu > 0.f)) return false;
    const float ln_nu = E.ln_nu_obs + lnf((float)Ecamd * kdotu);
    const float u_ph = ut * (g_tph + g_phph * Om);
    const float kb = lam - kdotu * u_ph;
    const float bb = g_phph + u_ph * u_ph;
    const float cosb2 = kb * kb * rcpf(kdotu * kdotu * bb);   // cos^2 of the pitch angle
    const float sinb2 = fmaxf(1.f - cosb2, 1e-6f);
    f.ln_th = ln_th;
    f.ln_nu = ln_nu;
    f.ln_sinb = 0.5f * lnf(sinb2);
    f.ln_g = E.ln_nu_obs - ln_nu;
    f.ln_dl = lnf(kdotu * Sigma * (float)hd) + E.ln_rg;   // k.u Sigma h: within FP32 range inside the flow
    q.lr = lr;
    q.lnvert = lnvert;
    q.soft = soft;
    q.R = R;
    q.z = z;
    return true;
}

template <int RADIO>
__device__ __forceinline__ bool flow_finish(const TraceParams& P, const EmisParams& E, double rd, double phid,
                                            double td, FlowSample& f, const FlowPending& q, const TurbCache& C) {
    float dens = 1.f;   // turbulence factor x horizon softening: one logarithm for both
    if (P.turb_on) {
        if constexpr (RADIO == bhd::RADIO_SHEARED) dens = sheared_factor_f32(P, (float)rd, q.R, q.z, phid, td);
        else dens = turb_blend_f32(q.tf, C);
    }
    const float ln_ne = E.ln_n0 - E.n_index * q.lr + q.lnvert + lnf(dens * q.soft);
    if (ln_ne < -69.07755f) return false;   // n_e < 1e-30
    const float ln_B = 0.5f * (E.ln_bconst + ln_ne - q.lr);
    f.ln_ne = ln_ne;
    f.ln_nuc = E.ln_nuc_per_B + ln_B;
    return true;
}

// dI = g^3 B_nu (1 - e^-tau), dtau = alpha dl (trace.cuh:131-143 with emission.cuh:167-174)
__device__ __forceinline__ void transfer(const EmisParams& E, const FlowSample& f, float ln_a, float ln_bnu,
                                         double* dI, double* dtau, float* etau) {
    const float tau = expf_(ln_a + f.ln_dl);
    const float e = expf_(-tau);   // also the step's transmission factor
    const float om = tau < 1e-4f ? tau * (1.f - 0.5f * tau) : 1.f - e;
    *dtau = (double)tau;
    *etau = e;
    *dI = (double)(expf_(3.f * f.ln_g + ln_bnu) * om);
}

// ---------------------------------------------------------------------------------------------
// Trilinear lookup of the baked surrogate at surrogate coordinates (ln X, ln Theta, ln sin theta_B),
// in two parts so the caller can evaluate the exact formula while the table's loads are in flight:
// the fetch computes the cell and issues the loads of its trust flag and eight corners (one memory
// round trip; the indices are in range whatever the flag says); the blend interpolates them. Outside
// the table, or in a cell the exact-routine judge flagged, the exact value stands. ln X above the
// table is clamped to its edge, where both coefficients sit at their floors.
struct SurrogateFetch {
    float2 v000, v100, v010, v110, v001, v101, v011, v111;
    float f0, f1, f2;
    bool use;   // inside the table and its cell not flagged (valid after the loads)
};

__device__ __forceinline__ void surrogate_fetch(const bhmlp::SurrogateTable& T, float lx, float lth, float ls,
                                                SurrogateFetch& F) {
    const float g0 = (fminf(lx, T.hi0) - T.lo0) * T.inv_d0, g1 = (lth - T.lo1) * T.inv_d1, g2 = (ls - T.lo2) * T.inv_d2;
    const bool in = g0 >= 0.f && g1 >= 0.f && g2 >= 0.f && g0 <= (float)(T.n0 - 1) && g1 <= (float)(T.n1 - 1) &&
                    g2 <= (float)(T.n2 - 1);
    // clamped in float before the conversion, so any input (NaN included) gives a cell in range;
    // inside the box this is the cell of the point (outside, its values are not used)
    const int i0 = (int)fminf(fmaxf(g0, 0.f), (float)(T.n0 - 2)), i1 = (int)fminf(fmaxf(g1, 0.f), (float)(T.n1 - 2)),
              i2 = (int)fminf(fmaxf(g2, 0.f), (float)(T.n2 - 2));
    // 32-bit offsets: bh_mlp_bake keeps the table at most 2^24 nodes
    const unsigned int s1 = (unsigned int)T.n0, s2 = (unsigned int)(T.n0 * T.n1);
    const unsigned char flag = T.bad[((unsigned int)i2 * (unsigned int)(T.n1 - 1) + (unsigned int)i1) *
                                         (unsigned int)(T.n0 - 1) + (unsigned int)i0];
    const float2* b = (const float2*)T.tab + ((unsigned int)i2 * s2 + (unsigned int)i1 * s1 + (unsigned int)i0);
    F.v000 = b[0]; F.v100 = b[1]; F.v010 = b[s1]; F.v110 = b[s1 + 1];
    F.v001 = b[s2]; F.v101 = b[s2 + 1]; F.v011 = b[s2 + s1]; F.v111 = b[s2 + s1 + 1];
    F.f0 = g0 - (float)i0;
    F.f1 = g1 - (float)i1;
    F.f2 = g2 - (float)i2;
    F.use = in && !flag;
}

__device__ __forceinline__ void surrogate_blend(const SurrogateFetch& F, float* fj, float* fa) {
    const float f0 = F.f0, f1 = F.f1, f2 = F.f2;
    float x00 = F.v000.x + (F.v100.x - F.v000.x) * f0, x10 = F.v010.x + (F.v110.x - F.v010.x) * f0;
    float x01 = F.v001.x + (F.v101.x - F.v001.x) * f0, x11 = F.v011.x + (F.v111.x - F.v011.x) * f0;
    float y00 = F.v000.y + (F.v100.y - F.v000.y) * f0, y10 = F.v010.y + (F.v110.y - F.v010.y) * f0;
    float y01 = F.v001.y + (F.v101.y - F.v001.y) * f0, y11 = F.v011.y + (F.v111.y - F.v011.y) * f0;
    float x0 = x00 + (x10 - x00) * f1, x1 = x01 + (x11 - x01) * f1;
    float y0 = y00 + (y10 - y00) * f1, y1 = y01 + (y11 - y01) * f1;
    *fj = x0 + (x1 - x0) * f2;
    *fa = y0 + (y1 - y0) * f2;
}

struct MixedParams {
    double r_far;     // far field (tolerance tol_far) at r >= r_far
    double r_f32;     // FP32 stages at r >= r_f32 (>= r_far; infinite: FP64 stages everywhere)
    double r_tol;     // scene tolerance below r_tol, tol_mid between r_tol and r_far
    double r_cut;     // emission, absorption and the transfer step cap only below r_cut
    double tol_mid;   // tolerance between r_tol and r_far
    double tol_far;   // tolerance of the far-field steps
    double inv_tol_t; // 1 / (absolute bound on a step's coordinate-time error); 0: no bound
    int need_t;       // integrate the coordinate time t (emission read at its time, or ray records)
    double tol_relax; // emission-region tolerance of a ray after relax_after attempts
    double relax_cap; // and the factor on its transfer step cap
    int relax_after;
    // reciprocals of the scene tolerance, tol_mid, tol_far and tol_relax (the error norm's 1 / tol)
    double inv_tol, inv_tol_mid, inv_tol_far, inv_tol_relax;
    EmisParams E;
    int use_surrogate;
    float ln_e2_over_c, ln_e2_over_mec;   // ln(e^2 / c), ln(e^2 / (m_e c))
    bhmlp::SurrogateTable S;
};

// Per-ray state of the mixed-policy tracer.
struct RayStateM {
    RayConst c;
    State s;
    double h, acc0, acc1, acc2, acc3, trans;
    float psi, s_sin;   // swept angle; sin theta of the current state
    TurbCache tc;       // the turbulence corners of the ray's last flow sample
    int ncross, accepted, flow_steps, it, fate, fp64_attempts;
    bool done, k1_f32;
};

template <int RADIO>
__device__ __forceinline__ void ray_init_m(const TraceParams& P, const MixedParams& M, double x, double yv,
                                           RayStateM& R) {
    bhd::camera_photon(P, x, yv, R.c, R.s);
    R.k1_f32 = R.s.y0 >= M.r_f32;
    if (R.k1_f32) {
        const RConst<float> cf = make_const<float>(P, R.c);
        Rhs<float> k = rhs<float>(cf, (float)R.s.y0, (float)R.s.y2);
        R.s.k1 = k.d1; R.s.k3 = k.d3; R.s.k4 = k.d4; R.s.k5 = k.d5;
    } else {
        const RConst<double> cd = make_const<double>(P, R.c);
        Rhs<double> k = rhs<double>(cd, R.s.y0, R.s.y2);
        R.s.k1 = k.d1; R.s.k3 = k.d3; R.s.k4 = k.d4; R.s.k5 = k.d5;
    }
    R.h = bhd::initial_step(R.s);
    R.acc0 = R.acc1 = R.acc2 = R.acc3 = 0.0;
    R.trans = 1.0;
    R.psi = 0.f;
    R.s_sin = sqrtf_(fmaxf((float)(1.0 - R.s.y2 * R.s.y2), 0.f));
    R.ncross = R.accepted = R.flow_steps = R.it = R.fate = R.fp64_attempts = 0;
    R.tc.v = Corners{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    R.tc.key = ~0u;   // grids hold fewer than 2^31 cells: no cell has this offset
    R.done = P.max_steps <= 0;
}

// One attempt of trace.cuh:81-163 under the mixed policy. TCACHE keeps the ray's turbulence cell
// in its state across attempts (the persistent engine); without it every flow sample loads its
// corners (the frame kernel, already at the register limit, would spill the cache in its loop).
template <int RADIO, bool TCACHE = true>
__device__ __forceinline__ void ray_step_m(const TraceParams& P, const MixedParams& M, RayStateM& R) {
    const bool flow = RADIO && P.flow_on;
    State n;
    double hc = R.h;
    if (flow && R.s.y0 < M.r_cut) {
        // the transfer step cap (trace.cuh:87-90) in FP32: a bound on the step, not part of its accuracy
        const float frac = (float)(R.it >= M.relax_after ? P.step_frac * M.relax_cap : P.step_frac);
        const float cap = frac * fminf((float)R.s.y0 * rcpf(fabsf((float)R.s.y1) + 1e-30f),
                                       fmaxf(R.s_sin, 0.05f) * rcpf(fabsf((float)R.s.y3) + 1e-30f));
        hc = fmin(hc, (double)cap);
    }
    double err;
    if (R.s.y0 >= M.r_f32) {
        const RConst<float> cf = make_const<float>(P, R.c);
        err = dp54<float>(cf, R.s, hc, M.inv_tol_far, M.inv_tol_t, n);
    } else {
        const RConst<double> cd = make_const<double>(P, R.c);
        if (R.k1_f32) {   // entering the FP64 region: FP64 derivative at the current state
            Rhs<double> k = rhs<double>(cd, R.s.y0, R.s.y2);
            R.s.k1 = k.d1; R.s.k3 = k.d3; R.s.k4 = k.d4; R.s.k5 = k.d5;
        }
        // inside the emission region the scene tolerance, relaxed for a ray that has already taken
        // relax_after attempts (an orbit near the photon sphere)
        const double inv_tol = R.s.y0 < M.r_tol ? (R.it >= M.relax_after ? M.inv_tol_relax : M.inv_tol)
                                                : (R.s.y0 < M.r_far ? M.inv_tol_mid : M.inv_tol_far);
        // coordinate time only where something reads it (wave-uniform: the frame's policy)
        err = M.need_t ? dp54<double, true>(cd, R.s, hc, inv_tol, M.inv_tol_t, n)
                       : dp54<double, false>(cd, R.s, hc, inv_tol, M.inv_tol_t, n);
        ++R.fp64_attempts;
    }
    R.k1_f32 = R.s.y0 >= M.r_f32;
    ++R.it;
    if (!(err <= 1.0)) {
        R.h = hc * fmax(0.2, 0.9 * pow_m02_d(err));
        if (!(err == err)) R.h = hc * 0.1;
        if (R.it >= P.max_steps) R.done = true;
        return;
    }
    ++R.accepted;
    // the flow sample's density-free part first: its turbulence loads are in flight during the ring
    // angle below
    const bool in_flow = flow && n.y0 < M.r_cut;
    FlowSample f;
    FlowPending q;
    TurbCache tl;
    TurbCache& tc = TCACHE ? R.tc : tl;
    bool emits = false;
    if (in_flow) emits = flow_begin<RADIO, TCACHE>(P, M.E, n.y0, n.y2, n.y4, n.y5, R.c.lam, R.c.E, hc, f, q, tc);
    // sin theta of the new state (FP64 1 - u^2, then FP32): the ring angle and the next step cap
    const float s1 = sqrtf_(fmaxf((float)(1.0 - n.y2 * n.y2), 0.f));
    if (RADIO) {
        // angle between the radial directions p = (s0 cos phi0, s0 sin phi0, u0) of the previous and
        // q of the new accepted state, from their chord without cancellation:
        // |q - p|^2 = (u1 - u0)^2 + (s1 - s0)^2 + 4 s0 s1 sin^2(dphi / 2), angle = 2 asin(|q - p| / 2),
        // with dphi taken in FP64 (no large-angle FP32 azimuth)
        const float s0 = R.s_sin, du = (float)n.y2 - (float)R.s.y2, ds = s1 - s0;
        const float hs = __sinf(0.5f * (float)(n.y4 - R.s.y4));
        const float chord = sqrtf_(du * du + ds * ds + 4.f * s0 * s1 * hs * hs);
        R.psi += 2.f * asin_pos(fminf(0.5f * chord, 1.f));
    }
    R.s_sin = s1;
    if ((R.s.y2 > 0.0) != (n.y2 > 0.0)) ++R.ncross;
    if (in_flow) {
        ++R.flow_steps;
        if (emits && flow_finish<RADIO>(P, M.E, n.y0, n.y4, n.y5, f, q, tc)) {
            float ln_j, ln_a, ln_bnu;
            if (M.use_surrogate) {
                // the table's loads first; the exact formula runs while they are in flight, and the
                // table replaces its values wherever the cell is trusted
                SurrogateFetch sf;
                surrogate_fetch(M.S, f.ln_nu - f.ln_nuc - (float)bhmlp::LN_2_9 - 2.f * f.ln_th - f.ln_sinb, f.ln_th,
                                f.ln_sinb, sf);
                synchrotron_exact(M.E, f.ln_ne, f.ln_th, f.ln_k2, f.ln_nu, f.ln_nuc, f.ln_sinb, &ln_j, &ln_a, &ln_bnu);
                if (sf.use) {
                    float fj, fa;
                    surrogate_blend(sf, &fj, &fa);
                    ln_j = f.ln_ne + M.ln_e2_over_c + f.ln_nuc + fj;
                    ln_a = f.ln_ne + M.ln_e2_over_mec - f.ln_nuc + fa;
                    ln_bnu = ln_j - ln_a;   // both within the table's floors: no cancellation
                }
            } else {
                synchrotron_exact(M.E, f.ln_ne, f.ln_th, f.ln_k2, f.ln_nu, f.ln_nuc, f.ln_sinb, &ln_j, &ln_a, &ln_bnu);
            }
            double dI, dtau;
            float etau;
            transfer(M.E, f, ln_a, ln_bnu, &dI, &dtau, &etau);
            double cI = R.trans * dI;
            int ord = (int)(R.psi * 0.31830988618379067f);
            R.acc0 += cI;
            // value selects, not a branch chain: the chain becomes a dynamically indexed stack array
            R.acc1 += ord == 1 ? cI : 0.0;
            R.acc2 += ord == 2 ? cI : 0.0;
            R.acc3 += ord >= 3 ? cI : 0.0;
            R.trans *= (double)etau;
            if (R.trans < 1e-5) {
                R.fate = -1;
                R.done = true;
                return;
            }
        }
    }
    const double r0 = R.s.y0;
    R.s = n;
    project_radial_m(P, R.c, R.s);
    double fac = fmin(5.0, fmax(0.2, 0.9 * pow_m02_d(err)));
    // Outside the scene-tolerance region the step follows the radial scale: there the Mino-time
    // step that meets the tolerance goes as 1/r (dr/dlambda ~ r^2), and the error-only controller
    // lags one step behind it - on the way out every other step was rejected.
    if (r0 >= M.r_tol) fac *= r0 * rcpd(R.s.y0);
    R.h = hc * fac;
    if (R.s.y0 < P.r_stop) {
        R.fate = -1;
        R.done = true;
        return;
    }
    if (R.s.y0 > P.r_escape && R.s.y1 > 0.0) {
        R.fate = 1;
        R.done = true;
        return;
    }
    if (R.it >= P.max_steps) R.done = true;
}

// Outputs of a finished ray; returns the number of FP64 step attempts.
__device__ __forceinline__ int ray_result_m(const TraceParams& P, const RayStateM& R, RayResult& out) {
    out.y[0] = R.s.y0; out.y[1] = R.s.y1; out.y[2] = R.s.y2; out.y[3] = R.s.y3; out.y[4] = R.s.y4;
    out.y[5] = R.s.y5;
    out.th_inf = 0.0;
    out.ph_inf = 0.0;
    if (R.fate == 1) {
        double sq = sqrt(fmax(1.0 - R.s.y2 * R.s.y2, 1e-30));
        double tq = (double)acosf(fmaxf(-1.f, fminf(1.f, (float)R.s.y2)));
        // acos in FP32 is not enough near the poles: one FP64 Newton step on cos
        tq = tq + (cos(tq) - R.s.y2) * rcpd(fmax(sin(tq), 1e-30));
        double dthdr = (-R.s.y3 * rcpd(sq)) * rcpd(R.s.y1), dphdr = R.s.k4 * rcpd(R.s.y1);
        out.th_inf = tq + R.s.y0 * dthdr;
        out.ph_inf = R.s.y4 + R.s.y0 * dphdr;
    }
    out.I[0] = R.acc0; out.I[1] = R.acc1; out.I[2] = R.acc2; out.I[3] = R.acc3;
    out.trans = R.trans;
    out.psi = (double)R.psi;
    out.fate = R.fate;
    out.ncross = R.ncross;
    out.order = (int)(R.psi * 0.31830988618379067f);
    out.steps = R.it;
    out.accepted = R.accepted;
    out.flow_steps = R.flow_steps;
    bhd::sky_result(P, R.trans, out);
    return R.fp64_attempts;
}

/* blackholes/src/trace.cuh:62-165 under the mixed policy. Returns the number of FP64 step attempts
   (near field and flow); out.steps counts all attempts. */
template <int RADIO>
__device__ __forceinline__ int trace_ray_mixed(const TraceParams& P, const MixedParams& M, double x, double yv,
                                               RayResult& out) {
    RayStateM R;
    ray_init_m<RADIO>(P, M, x, yv, R);
    while (!R.done) ray_step_m<RADIO, false>(P, M, R);
    return ray_result_m(P, R, out);
}

#pragma clang fp contract(off)

}  // namespace bhm32
