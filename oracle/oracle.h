/* oracle.h - CPU FP64 restatement of bhrt's per-pixel Kerr tracer. TEST INFRASTRUCTURE ONLY.

   Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
   only as the checker (or as the timed CPU baseline). The product never links or calls it.

   Built twice from oracle.cpp (see oracle/Makefile):
   - liboracle_libm.so : transcendentals from glibc, exactly like the reference's host build. Used
     to pin the restatement against the reference compiled from its own sources (oracle/_ref).
   - liboracle_bhm.so  : transcendentals from spacetime-optics_amd/csrc/bh_math.h. Used as the
     bit-exact checker of the GPU tracer, which uses the same elementary functions.

   Layouts mirror the reference's POD structs (blackholes/src/scene.hpp:12-80) for the fields on the
   radio/vacuum path; table pointers are plain host arrays. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Parameters of one bake of the sheared turbulence (sheared.cpp; render.cu:615-619)
typedef struct {
    double a, r_isco, cam_t, cam_r;
    float lr0, lr1, zmax, d0, dd, amp, scale, tcorr;
    uint32_t seed;   // scene seed ^ 0x51f15eed
    int nphi, nlr, nz, nt;
} or_sheared;

typedef struct {
    double a, rh, r_isco;
    double tol;
    int max_steps;
    double r_escape, r_stop, step_frac;
    int mode;   // 0 visible (vacuum geodesics only: disk and sky are out of scope), 1 radio
    int flow_on;
    double n0, n_index, theta_index, theta0, h_r, beta_p, r_flow_max, nu_obs, rg_cm, flow_kep;
    float flow_turb_amp;
    // Turbulence: density factor on a cell-centred n^3 Cartesian grid spanning [-half, half]^3,
    // x fastest. n = 0 or grid == NULL disables it. With turb_nt > 1 the grid holds turb_nt such
    // slices at coordinate times turb_t0 + k / turb_inv_dt, blended linearly in time.
    const float* turb_grid;
    int turb_n;
    double turb_half, turb_inv_dx;
    int turb_nt;
    double turb_t0, turb_inv_dt;
    const float* k2_lntab;
    int k2_n;
    float k2_l0, k2_dl;
    uint32_t seed;
    float flow_turb_scale, flow_tcorr;
    // Sheared turbulence (turb_model 1): slices baked by or_sheared_bake with parameters ts
    int turb_model;
    const float* ts_slices;
    or_sheared ts;
    // Visible mode: procedural checkerboard celestial sphere on the galactic (u, v) of sky_uv
    // (sky_nu x sky_nv squares, bright sky_scale, dark 0.1 sky_scale); sky_rot: black-hole frame ->
    // galactic, row-major (scene_build.cpp:181-204)
    int sky_on, sky_nu, sky_nv;
    float sky_scale;
    double sky_rot[9];
} or_scene;

typedef struct {
    double r, th, ph, t;
    double kp[11];   // KerrPoint: r, s, c, Sigma, Delta, A, omega, e_nu, e_psi, sqrtSigma, sqrtSigmaDelta
    double F[3], R[3], U[3], beta[3];
    double cx, cy, pix;
    int W, H, spp;
    uint32_t seed;
} or_frame;

typedef struct {
    double y[6];          // state when the ray stopped: r, dr/dtau, u = cos(theta), du/dtau, phi, t
    double th_inf, ph_inf;
    double I[4];          // radio: I_nu total, n = 1, n = 2, n >= 3 (FP64 accumulation)
    double trans;
    double psi;
    int fate;             // +1 escaped, -1 captured or opaque, 0 step limit
    int ncross;           // equatorial crossings (winding count)
    int order;            // floor(psi / pi)
    int steps;            // DP54 step attempts
    int accepted;
    int flow_steps;       // accepted steps inside r < r_flow_max (radio)
    float sky_u, sky_v;   // visible mode, escaped rays: galactic texture coordinates (trace.cuh:25-33)
} or_ray_out;

// kerr.hpp
void or_kerr_point(double a, double r, double th, double out[11]);
double or_horizon_radius(double a);
double or_isco_radius(double a);
void or_camera_photon(double a, const double kp[11], double phi0, double t0, const double d[3],
                      const double beta[3], double out[9]);
double or_radial_potential(double a, double lam, double eta, double r);
double or_polar_potential(double a, double lam, double eta, double u);
void or_geodesic_rhs(double a, double lam, double eta, const double* y, double* d);
double or_dp54_step(double a, double lam, double eta, const double* y, const double* k1, double h,
                    double tol, double* yout, double* k7);
void or_project_radial(double a, double lam, double eta, double* y, double* k);
double or_initial_step(const double* y);
int or_vacuum_fate(double a, const double photon[9], double tol, double r_escape, int max_steps);
int or_critical_curve_point(double a, double theta_o, double rp, double* alpha, double* beta_plus);

// shot.cpp
void or_make_frame(double a, double r, double incl, double phi, double roll, double yaw, double pitch,
                   double fov, int fov_in_m, double zoom, double target, double target_x,
                   double target_y, double off_x, double off_y, double time_m, int orbit, int W, int H,
                   uint32_t seed, or_frame* F);
// Bisection on the vacuum fate along image direction psi from the frame's centre (cx, cy ignored).
double or_find_critical_point(double a, const or_frame* F, double psi);

// device_util.cuh / render.cu sampling
uint32_t or_pcg(uint32_t v);
void or_pixel_seed(uint32_t frame_seed, int p, float* r0, float* r1);
void or_r2_sample(uint32_t n, float r0, float r1, float* x, float* y);
float or_table_lerp(const float* t, int n, float x0, float dx, float x);
// Image-plane offsets of sample s of pixel (px, py); jitter = 0 -> pixel centre.
void or_sample_offsets(const or_frame* F, int px, int py, int s, int jitter, double* tx, double* ty);

// scene_build.cpp K2 table: ln K2(1/Theta) on ln Theta in [ln 0.02, ln 1e4]
void or_k2_table(int n, float* tab, float* l0, float* dl);

// emission.cuh / trace.cuh
double or_turb_factor(const or_scene* S, double r, double u, double phi, double t);
void or_flow_step(const or_scene* S, double r, double u, double phi, double t, double lam, double Ecam,
                  double h, double* dI, double* dtau);
void or_escape_angles(const double* y, const double* k, double* th, double* ph);
// Galactic texture coordinates of an escape direction (trace.cuh:25-33) through the row-major rotation rot
void or_sky_uv(const double rot[9], double th, double ph, float* u, float* v);
// Procedural checkerboard radiance at (u, v): nu x nv squares, bright `scale`, dark 0.1 scale
float or_sky_checker(int nu, int nv, float scale, float u, float v);
void or_trace_ray(const or_scene* S, const or_frame* F, double x, double yv, or_ray_out* out);
// Sample s of pixel (px, py): or_sample_offsets then or_trace_ray
void or_trace_sample(const or_scene* S, const or_frame* F, int px, int py, int s, int jitter, or_ray_out* out);

// Frame drivers (std::thread, dynamic row scheduling). Rays: pixel list (px[i], py[i]), spp samples
// each (jitter 0 = centre rays). out_px (nullable): per pixel mean of I[0..3] over the samples.
// out_rays (nullable, spp = 1 only): per-ray records. Returns the number of threads used.
int or_trace_pixels(const or_scene* S, const or_frame* F, const int* px, const int* py, int n, int spp,
                    int jitter, int nthreads, double* out_px, or_ray_out* out_rays);

// Adaptive sampling of a W x H frame (F->W, F->H), restating render.cu:89-242 and 677-721 with this
// project's deterministic sums:
//  - pass 1: samples 0..spp-1 of every pixel; per pixel the FP64 sum of I[0..3] in sample order and
//    the float luminance sums ys += Y, yq += Y*Y with Y = (float)I[0] (radio) or (float)I[1] (visible);
//  - yref = (float)(0.03 * fsum / (n * spp) + 1e-37), fsum = FP64 sum of the ys in blocks of 256
//    consecutive pixels (each block in pixel order, then the blocks in order);
//  - extra samples per pixel as k_adapt (float arithmetic, render.cu:130-139), dilated by the 3 x 3
//    maximum and rounded up to chunks of 8 (render.cu:141-150, at most spp_max - spp);
//  - pass 2: samples spp .. spp + 8 * chunks - 1 of each pixel, added to the pass-1 sum in order;
//  - pixel value = FP64 sum / sample count.
// out_px: W*H*4 doubles; out_nsamp (nullable): W*H sample counts. Returns the total sample count.
long long or_render_adaptive(const or_scene* S, const or_frame* F, int spp, int spp_max, float target_err,
                             int jitter, int nthreads, double* out_px, uint32_t* out_nsamp);

// Elementary function fn (numbering of bh_math_eval in include/bhrt.h) on n inputs, with this
// build's functions (glibc or bh_math.h); 9 = sqrt, 10 = x / y.
void or_math_eval(int fn, const double* x, const double* y, double* out, int n);

// Self-test restatement (main.cpp:323-372): fills 8 rows of (a, incl, traced -, traced +,
// analytic -, analytic +) and the unprojected radial-constraint drift.
void or_selftest(double rows[8][6], double* drift, int* drift_steps);

/* Sheared turbulence (sheared.cpp; emission.cuh:20-51,113-122, device_util.cuh:15-71,
   render.cu:449-460,615-619). dims: phi, ln R, z/R cells and slices (NULL or 0 -> 512, 192, 48, 24).
   or_sheared_bake fills nt slices of nz x nlr x nphi floats (phi fastest). */
void or_sheared_params(const or_scene* S, double cam_t, double cam_r, const int dims[4], or_sheared* B);
float or_sheared_gnoise(float x, float y, float z, uint32_t seed);
float or_sheared_fbm(float x, float y, float z, int octaves, uint32_t seed);
float or_sheared_noise(const or_sheared* B, double R, double phi, double z, double t);
float or_sheared_tex(const float* T, int nx, int ny, int nz, float x, float y, float z);
float or_sheared_cell(const or_sheared* B, int i, int j, int l, int k);
void or_sheared_bake(const or_sheared* B, float* out, int nthreads);
float or_sheared_factor(const or_scene* S, double r, double R, double z, double phi, double t);

/* Radio post step and tone map (post.cpp; render.cu:244-407,742-791): beam blur in place on px
   (W*H*4 floats), the displayed-intensity statistics, the radio tone map to W*H*3 bytes. */
void or_post_blur(float* px, int W, int H, double sigma);
void or_frame_stats(const float* px, int W, int H, float hide_direct, double percentile, double* lum_percentile,
                    double* lum_max);
void or_tonemap_radio(const float* px, int W, int H, int colormap, float exposure, float gamma, float overlay,
                      float hide_direct, float fade, float vignette, float grain, uint32_t seed, uint8_t* rgb);
struct OrGlyph;
struct OrRect;
void or_hud_draw(uint8_t* rgb, int W, int H, const struct OrGlyph* glyphs, int ng, const struct OrRect* rects, int nr,
                 const uint8_t* atlas, int cell_w, int cell_h, float off);

#ifdef __cplusplus
}
#endif
