/* bhrt.h - C ABI of the MI355X (gfx950) Kerr tracer: the drop-in replacement for bhrt's GPU
   Renderer on the per-pixel geodesic + thermal-synchrotron path.

   Reference interface replaced (paths relative to the reference repository root):
     class Renderer                       blackholes/src/renderer.hpp:72-88
     SceneGPU / FrameGPU PODs             blackholes/src/scene.hpp:12-80
     SceneTables / RenderSettings         blackholes/src/renderer.hpp:19-29
     RenderStats                          blackholes/src/renderer.hpp:64-68
     HdrFrame.px (W*H*4 floats)           blackholes/src/renderer.hpp:37-46

   Conventions: plain C types only; integer status codes (BH_OK == 0), never abort; error text is
   copied to the caller's buffer or read with bh_last_error(). One context per device; a context is
   not reentrant (callers serialise). Every call is synchronous, like the reference's (render.cu:697,
   713, 725, 770), except bh_render_resident, which leaves the frame in device memory. */
#ifndef BHRT_H
#define BHRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_ERR_ARG (-1)
#define BH_ERR_HIP (-2)
#define BH_ERR_NODEVICE (-3)
#define BH_ERR_UNSUPPORTED (-4)

enum { BH_MODE_VISIBLE = 0, BH_MODE_RADIO = 1 };      /* scene.hpp:10 RenderMode */
enum { BH_PREC_FP64 = 0, BH_PREC_MIXED = 1 };         /* plain FP64 (parity) or mixed FP32/FP64 */
enum { BH_EMIS_EXACT = 0, BH_EMIS_MLP = 1 };          /* synchrotron coefficients: exact or surrogate */
enum { BH_ENGINE_AUTO = 0, BH_ENGINE_FRAME = 1 };     /* kernel behind bh_render (bh_trace_opts.engine) */
enum { BH_FATE_LIMIT = 0, BH_FATE_ESCAPED = 1, BH_FATE_CAPTURED = -1 };
enum { BH_ORDER_ROWS = 1, BH_ORDER_IMPACT = 2, BH_ORDER_CRITICAL = 3 };   /* tile order of a persistent launch's queues */

/* SceneGPU (scene.hpp:12-69): the scalars on the vacuum/radio path, and the celestial sphere of the
   visible mode. The thin disk and colour fields are out of scope (visible-mode presentation) and
   absent. The reference samples a star-map texture (sky_tex, scene.hpp:58-61); this path's sky is a
   procedural checkerboard on the same galactic texture coordinates (trace.cuh:25-33). */
typedef struct bh_scene {
    double a, rh, r_isco;
    double tol;             /* DP54 tolerance (render.tolerance) */
    int max_steps;
    double r_escape, r_stop, step_frac;
    int mode;               /* BH_MODE_* ; visible = vacuum geodesics only */
    int flow_on;
    double n0, n_index, theta_index, theta0, h_r, beta_p, r_flow_max, nu_obs, rg_cm, flow_kep;
    float flow_turb_amp;    /* > 0 enables the turbulence (bh_tables.turb_model) */
    uint32_t seed;
    float flow_turb_scale;  /* sheared turbulence: feature size in units of R (flow.turbulence_scale) */
    float flow_tcorr;       /* sheared turbulence: correlation time in orbital periods (flow.correlation_orbits) */
    int sky_on;             /* visible mode: escaped rays take the checkerboard sky (sky_on, scene.hpp:58) */
    int sky_nu, sky_nv;     /* squares of the checkerboard along galactic u (longitude) and v (latitude) */
    float sky_scale;        /* radiance of the bright squares; dark ones are 0.1 of it (sky.brightness) */
    double sky_rot[9];      /* black-hole frame -> galactic, row-major (sky_rot, scene.hpp:61) */
} bh_scene;

/* SceneTables (renderer.hpp:19-23) plus the turbulence input. The K2 table is ln K2(1/Theta) on a
   ln Theta grid (scene_build.cpp:159-170). Two turbulence models (density factor of the flow):
   - BH_TURB_SHEARED, the reference's: sheared, finite-lifetime fBm noise (emission.cuh:30-51)
     baked by the library on the device for each frame's camera time into ts_dims[3] slices of
     retarded time, each ts_dims[2] x ts_dims[1] x ts_dims[0] cells over (z/R, ln R, phi)
     (render.cu:447-460,594-619; 0 -> the reference's 512 x 192 x 48 x 24), read with trilinear
     filtering (wrap in phi, clamp elsewhere, weights rounded to 1/256 like a texture unit) and
     interpolated between slices (emission.cuh:112-123). turb is ignored.
   - BH_TURB_GRID: turb_nt time slices of a cell-centred turb_n^3 Cartesian grid over
     [-turb_half, turb_half]^3 (x fastest) given by the caller. With turb_nt > 1, slice k sits at
     coordinate time turb_t0 + k*turb_dt and the tracer interpolates linearly in the emission time
     of each step (slow light). */
enum { BH_TURB_GRID = 0, BH_TURB_SHEARED = 1 };
typedef struct bh_tables {
    const float* k2_ln;
    int k2_n;
    float k2_l0, k2_dl;
    const float* turb;
    int turb_n, turb_nt;
    double turb_half, turb_t0, turb_dt;
    int turb_model;         /* BH_TURB_* */
    int ts_dims[4];         /* sheared: phi, ln R, z/R cells and time slices (0 -> 512, 192, 48, 24) */
} bh_tables;

/* FrameGPU (scene.hpp:71-80). kp = KerrPoint at the camera (kerr.hpp:20-25): r, sin th, cos th,
   Sigma, Delta, A, omega, e_nu, e_psi, sqrt(Sigma), sqrt(Sigma Delta). */
typedef struct bh_frame {
    double r, th, ph, t;
    double kp[11];
    double F[3], R[3], U[3], beta[3];
    double cx, cy, pix;
    int W, H, spp;
    uint32_t seed;
} bh_frame;

/* RenderSettings (renderer.hpp:25-29) */
typedef struct bh_settings {
    int W, H, spp, spp_max;
    float target_err;
} bh_settings;

/* Per-call options (new; the reference has one fixed path). */
typedef struct bh_trace_opts {
    int precision;    /* BH_PREC_* */
    int emissivity;   /* BH_EMIS_* (radio only) */
    int jitter;       /* 0: pixel-centre samples; 1: R2 + PCG jitter as render.cu:70-85 */
    int spp;          /* samples per pixel; 0 -> frame->spp */
    double tol_far;   /* mixed: tolerance of the far-field steps (0 -> vacuum: max(1e-6, tol);
                         radio: 1e-4 * (pix * r) / 0.04 M, the pixel's size at the hole, at most 1e-4,
                         times sin(theta_cam) / sin(15 deg) for a camera within 15 degrees of the spin
                         axis, and at least max(1e-6, tol)) */
    double r_near;    /* mixed, vacuum: FP64 below this radius (0 -> max(3 r_+, r_ph) + 1) */
    double r_emit;    /* mixed, radio: the emission region is r < max(r_emit, r_near): scene tolerance,
                         emission, absorption and the transfer step cap there; tol_mid from there to
                         the flow's edge, tol_far beyond (0 -> the radius outside which the scene's
                         flow model, in its mid-plane, emits less than 1e-6 of its r^2-weighted
                         emissivity and has a radial optical depth below 1e-7) */
    int far_fp32;     /* mixed: 0 -> FP64 stages everywhere (default: on MI355X FP32 arithmetic without
                         packed FP32 runs at the FP64 rate, and a per-lane precision switch costs
                         both stage codes per wave); 1 -> FP32 stages in the far field (radio: outside
                         the flow; vacuum: outside r_near), same tolerances */
    int full_flow;    /* mixed, radio: 0 (default) -> the flow outside the emission region is traced as
                         vacuum with tol_mid = max(max(1e-6, tol), tol_far / 10); 1 -> emission,
                         absorption and the step cap over the whole flow as the FP64 tracer, with
                         tol_mid = max(1e-6, tol) */
    int engine;       /* bh_render / bh_render_resident / bh_trace_tile: BH_ENGINE_AUTO (0) -> the
                         persistent engine (bh_autotune's last result for this precision, or the
                         defaults): one sample per pixel as bh_trace_batch of one frame, more as the
                         first pass of bh_render_adaptive alone; frames beyond 2^31 samples, or whose
                         samples would take more than 1 GiB (32 bytes each), on the frame kernel.
                         BH_ENGINE_FRAME -> always the frame kernel. Both give the same pixels,
                         counters and ray records. */
    double tol_t;     /* mixed: absolute bound (M) on one step's coordinate-time error, which the
                         reference's error norm leaves out; it matters where the turbulence depends on
                         the emission time (grid time slices, the sheared model). 0 -> the default:
                         such scenes get a bound, others none; < 0 -> no bound */
    int relax_after;  /* mixed: after this many attempts a ray's emission-region tolerance becomes
                         tol_relax (rays that orbit near the photon sphere); 0 -> the default, < 0 -> never */
    double tol_relax; /* mixed: that tolerance (0 -> the default) */
    double relax_cap; /* mixed: and the factor on its transfer step cap (0 -> the default) */
} bh_trace_opts;

/* RenderStats (renderer.hpp:64-68) extended with the integrator counters. */
typedef struct bh_stats {
    double ms_trace;          /* host wall time of the call */
    double ms_kernel;         /* HIP-event time of the trace kernels on the context stream */
    long long samples;
    long long step_limited;   /* rays with fate 0 */
    long long step_attempts;  /* DP54 steps tried (accepted + rejected) */
    long long steps_accepted;
    long long flow_steps;     /* accepted steps inside the flow (radio) */
    long long fp64_steps;     /* mixed precision: attempts taken on the FP64 path */
    double flux_sum;          /* sum over samples of I_nu (radio) */
} bh_stats;

/* Per-ray record of the tile interface (new: trace.cuh:62 keeps these internal). */
typedef struct bh_ray_out {
    double y[6];              /* r, dr/dtau, u = cos th, du/dtau, phi, t when the ray stopped */
    double th_inf, ph_inf;    /* escape direction (trace.cuh:49-55), escaped rays only */
    double I[4];              /* radio: I_nu, and its n = 1, 2, >= 3 parts */
    double trans;
    double psi;               /* swept angle (radio) */
    int fate, ncross, order, steps, accepted, flow_steps;
    float sky_u, sky_v;       /* visible mode, escaped rays: galactic texture coordinates (uv_out, trace.cuh:62) */
} bh_ray_out;

typedef struct bh_ctx bh_ctx;

/* int bh_device_count()                                   renderer.hpp:70, render.cu:491-495 */
int bh_device_count(void);

/* Renderer::init(device, scene, tables, sky, settings, err)   renderer.hpp:76-77, render.cu:531-656
   The sky texture is out of scope. Tables are copied to device memory. */
int bh_create(int device, const bh_scene* scene, const bh_tables* tables, const bh_settings* settings,
              bh_ctx** out, char* err, size_t errlen);

/* Renderer::update_scene(scene): scalars only, tables stay    renderer.hpp:78, render.cu:658-675 */
int bh_update_scene(bh_ctx* ctx, const bh_scene* scene);

/* Current scene scalars of a context. */
int bh_get_scene(const bh_ctx* ctx, bh_scene* scene);

/* Renderer::render(frame, post, percentile, out, stats)      renderer.hpp:79, render.cu:755-795
   Frame interface: traces W*H pixels of `frame` (spp samples each, plain mean; adaptive sampling
   and post-processing are not part of this path) into px_out[W*H*4] = (I, I_n1, I_n2, I_n>=3) per
   pixel, row-major from the top-left. px_out may be NULL. */
int bh_render(bh_ctx* ctx, const bh_frame* frame, const bh_trace_opts* opts, float* px_out, bh_stats* stats);

/* As bh_render, but the frame stays in device memory (float4 per pixel, returned in *d_px, owned by
   the context and valid until the next call); only the counters come back. Used to time the tracer
   with its inputs and outputs resident in HBM. */
int bh_render_resident(bh_ctx* ctx, const bh_frame* frame, const bh_trace_opts* opts, const void** d_px,
                       bh_stats* stats);

/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc): bh_render_resident on the
   frame kernel (the BH_ENGINE_FRAME path, same pixels and counters), writing the frame into the
   caller's device buffer d_px (W*H float4 = W*H*4 floats) and allocating nothing. The frame kernel's
   per-workgroup partials live in a fixed device array of the library (3 MiB per device, held until
   the process ends), so frames are limited to 65536 workgroups of 16 x 16 pixels (4096 x 4096);
   larger frames fail with BH_ERR_ARG. Every call on one device shares that array: calls that overlap
   on one device (several contexts or threads) race and must be serialised by the caller. A context
   with sheared turbulence must have traced one frame (any entry point) before, so the slice buffers
   exist; else BH_ERR_UNSUPPORTED. Synchronous, like bh_render. */
int bh_render_noalloc(bh_ctx* ctx, const bh_frame* frame, const bh_trace_opts* opts, void* d_px, bh_stats* stats);

/* Tile interface (k_pass1's per-pixel launch, render.cu:89-128, with every per-ray output): rays of
   pixels x0 <= px < x0+w, y0 <= py < y0+h, one sample each (centre or jittered sample 0), written
   row-major to out[w*h]. opts->engine BH_ENGINE_AUTO (0): the persistent engine with bh_autotune's
   last launch for the precision (or the defaults); BH_ENGINE_FRAME: the frame kernel. Both give the
   same records. */
int bh_trace_tile(bh_ctx* ctx, const bh_frame* frame, int x0, int y0, int w, int h, const bh_trace_opts* opts,
                  bh_ray_out* out);

/* Renderer::radio_flux(frame, n, pix_sr): flux density (Jy) of an n x n, spp = 2 jittered frame
   renderer.hpp:81, render.cu:797-809 */
int bh_radio_flux(bh_ctx* ctx, const bh_frame* frame, int n, double pix_sr, double* jy);

/* vacuum_fate (kerr.hpp:184-203) for n photons; photons[9*i..] = (lam, eta, E, y[6]) as
   camera_photon returns them. fates[i] in {+1, -1, 0}. Used by the critical-curve search
   (shot.cpp:148-164) and the self-test. */
int bh_vacuum_fates(bh_ctx* ctx, const double* photons, int n, double tol, double r_escape, int max_steps,
                    int* fates);

/* Renderer::render's adaptive sampling (render.cu:677-721, k_pass1 / k_adapt / k_dilate / scan /
   k_pass2 / k_resolve, render.cu:89-242): settings.spp samples per pixel (opts->spp > 0 overrides),
   then, where the standard error of the pixel's luminance exceeds settings.target_err of (its mean +
   3% of the frame mean), more samples up to settings.spp_max, in chunks of 8, dilated over the 3 x 3
   neighbourhood (defaults 8 / 128 / 0.03 when the settings are 0). Samples continue each pixel's
   R2 + PCG sequence (jittered when opts->jitter). Sums are deterministic: a pixel is the FP64 sum of
   its samples in sample order over its sample count. nsamp_out (nullable): samples per pixel. */
int bh_render_adaptive(bh_ctx* ctx, const bh_frame* frame, const bh_trace_opts* opts, float* px_out,
                       unsigned int* nsamp_out, bh_stats* stats);

/* ---- Persistent batched tracing (new; the reference renders one frame per launch sequence,
   render.cu:677-730). Frames of different size and scene (e.g. spin) are traced by one persistent
   launch: resident workgroups pull 8 x 8 pixel tiles from per-frame work queues, retire finished
   rays every `interval` steps and refill free lanes. The launch sequence (queue reset, persistent
   kernel, statistics reduction) is captured once into a hipGraph and replayed. ---- */
typedef struct bh_batch_frame {
    bh_scene scene;     /* per-frame scene (tables and turbulence come from the context) */
    bh_frame frame;     /* W, H, camera; one sample per pixel (centre or jittered sample 0) */
} bh_batch_frame;

typedef struct bh_batch_cfg {
    int block;          /* threads per workgroup: 64, 128 or 256 (0 -> the default) */
    int grid_per_cu;    /* resident workgroups per compute unit (0 -> the default); capped at what the
                           CU holds at once for the kernel and block size */
    int interval;       /* integration steps between retire/refill (0 -> the default) */
    int refill_min;     /* free lanes needed to refill (0 -> the default) */
    int use_graph;      /* capture and replay the launch as a hipGraph */
    int waves;          /* 0: kernel compiled for the compiler's register budget; 3 or 4: compiled for at
                           least that many resident waves per SIMD (fewer registers) */
    int order;          /* tiles of each frame queued in BH_ORDER_ROWS (row) order, BH_ORDER_IMPACT
                           order (by the impact parameter of the tile's centre ray, smallest first) or
                           BH_ORDER_CRITICAL order (by the predicted length of the tile's centre ray,
                           longest first: -ln |min R(r) / r^4| over the spherical photon orbits' radii,
                           which grows with the number of orbits near the critical curve); 0 -> the
                           default */
    int prio_age;       /* > 0: a wave holding a ray that has taken prio_age steps raises its issue
                           priority; < 0: off; 0 -> the default. A NULL cfg (and no autotuned one)
                           means, for the mixed tracer, 128 / 4 / 32 / 32 / graph / 3 / critical order /
                           256, for the FP64 tracer 256 / 8 / 64 / 48 / graph / 3 / row order / 256. */
} bh_batch_cfg;

/* Traces every frame of `frames`; px_out (nullable) receives the frames one after the other,
   W*H*4 floats each. stats aggregates all frames. */
int bh_trace_batch(bh_ctx* ctx, const bh_batch_frame* frames, int n, const bh_trace_opts* opts,
                   const bh_batch_cfg* cfg, float* px_out, bh_stats* stats);

/* Times bh_trace_batch over a grid of configurations on the given workload and returns the fastest
   (best) with its kernel time (best_ms): the median of three timed launches after a warm-up one. tried
   (nullable, 8 doubles per configuration: block, grid_per_cu, interval, refill_min, ms, waves, order,
   prio_age) receives up to max_tried rows;
   *n_tried their number. */
int bh_autotune(bh_ctx* ctx, const bh_batch_frame* frames, int n, const bh_trace_opts* opts, bh_batch_cfg* best,
                double* best_ms, double* tried, int max_tried, int* n_tried);

/* ---- Thermal-synchrotron surrogate (new: the reference has only the exact formula,
   blackholes/src/emission.cuh:159-170). A 3 -> 64 -> 64 -> 64 -> 2 SiLU MLP of
   (log10 nu/nu_c, Theta_e, theta_B) -> (ln f_j, ln f_a), f_j = j / (n_e e^2 nu_c / c),
   f_a = alpha / (n_e e^2 / (m_e c nu_c)), trained on the GPU on samples of the exact routine. ---- */
typedef struct bh_mlp_train_cfg {
    int n_samples;      /* training samples (0 -> 2^22) */
    int batch;          /* samples per Adam step, multiple of 16 (0 -> 8192) */
    int epochs;         /* passes over the samples (0 -> 200) */
    float lr, lr_min;   /* cosine-decayed Adam learning rate (0 -> 2e-3, 2e-5) */
    float hot_weight;   /* extra loss weight where ln f_j > -12 (< 0 -> 20) */
    uint64_t seed;      /* sampler and initialisation seed (0 -> 1) */
} bh_mlp_train_cfg;

typedef struct bh_mlp_report {
    double ms_sample, ms_train;
    int steps;
    float loss_first, loss_last;
    double train_flop;
} bh_mlp_report;

typedef struct bh_mlp_bake_report {
    double ms_bake, ms_judge;
    long long cells, hot_cells, failed_hot_cells;
} bh_mlp_bake_report;

/* Samples the exact routine over the box and trains the surrogate on the context's device. */
int bh_mlp_train(bh_ctx* ctx, const bh_mlp_train_cfg* cfg, bh_mlp_report* report);
/* Parameters (8706 floats: W1[3][64] b1 W2[64][64] b2 W3[64][64] b3 W4[64][2] b4) and the
   normalisation (10 floats: input centres[3], input half-widths[3], output means[2], scales[2]).
   The network's inputs are the surrogate coordinates s = (ln X, ln Theta_e, ln sin theta_B),
   X = 10^u0 / ((2/9) Theta_e^2 sin theta_B), ln X clamped to at most 14, sin theta_B to at least
   1e-3; network input k = (s_k - centre_k) / half-width_k. */
int bh_mlp_get(bh_ctx* ctx, float* params, float* norm);
int bh_mlp_set(bh_ctx* ctx, const float* params, const float* norm);
/* Inference kernel on n raw inputs (log10 nu/nu_c, Theta_e, theta_B) -> (ln f_j, ln f_a). */
int bh_mlp_infer(bh_ctx* ctx, const float* in, int n, float* out);
/* The exact routine (FP64) at the same inputs -> (ln f_j, ln f_a). */
int bh_mlp_exact(bh_ctx* ctx, const float* in, int n, double* out);
/* Loss and gradient of the training kernels for given parameters, normalisation, normalised inputs
   xn[n][3], raw targets tn[n][2] and weights w[n] (n a multiple of 16). */
int bh_mlp_grad(bh_ctx* ctx, const float* params, const float* norm, const float* xn, const float* tn, const float* w,
                int n, float* loss, float* grad);
/* Tabulates the trained network on an n0 x n1 x n2 lattice over the surrogate coordinates
   (ln X in [-17, 14], ln Theta in [ln 0.5, ln 100], ln sin theta_B in [ln 1e-3, 0]; 0 -> 1024 x 128
   x 8) with the inference kernel and judges every cell against the exact routine at four interior
   points (|d ln f| > judge_tol, 0 -> 2e-3, fails it). The mixed tracer with BH_EMIS_MLP reads this
   table and uses the exact formula in failed cells and outside the box. */
int bh_mlp_bake(bh_ctx* ctx, int n0, int n1, int n2, float judge_tol, bh_mlp_bake_report* report);

/* The sheared turbulence slices for a frame's camera (bh_tables.turb_model = BH_TURB_SHEARED): bakes
   them as a trace of that frame would (render.cu:679-682) and copies them to out, which holds
   ts_dims[3] x ts_dims[2] x ts_dims[1] x ts_dims[0] floats (slice, z/R, ln R, phi; phi fastest);
   n = the number of floats out can hold. */
int bh_sheared_slices(bh_ctx* ctx, const bh_frame* frame, float* out, size_t n);

/* ---- Radio post-processing and display (the radio half of Renderer::render's post step,
   render.cu:742-753,756-792, and of Renderer::tonemap, render.cu:811-846, without the HUD overlay).
   Frames are linear HDR float4 pixels (I, I_n1, I_n2, I_n>=3) as bh_render writes them. ---- */
typedef struct bh_post_params {     /* PostParams (renderer.hpp:31-35), radio fields */
    double blur_sigma_px;           /* Gaussian beam sigma in pixels (applied when > 0.5) */
    float hide_direct;              /* fraction of the direct (n = 0) light left out of the statistics */
} bh_post_params;

/* Beam blur in place on px (W*H*4 floats, host memory), then the frame statistics of the displayed
   intensity I - hide_direct * max(I - I_n1 - I_n2 - I_n>=3, 0): lum_max, and lum_percentile = the
   `percentile` quantile of the pixels above 1e-5 of the maximum (Renderer::render,
   render.cu:756-792). Either output pointer may be NULL. */
int bh_post_radio(bh_ctx* ctx, float* px, int W, int H, const bh_post_params* params, double percentile,
                  double* lum_percentile, double* lum_max);

typedef struct bh_tone_params {     /* ToneParams (renderer.hpp:48-59), radio fields */
    int colormap;                   /* 0 afmhot (EHT), 1 inferno */
    float exposure;                 /* multiplies the displayed intensity */
    float gamma;                    /* display gamma */
    float overlay;                  /* photon-ring order tint, 0..1 */
    float hide_direct;              /* fraction of the n = 0 light removed */
    float fade;                     /* fade to black, 1 = none */
    float vignette;
    float grain;                    /* film grain amplitude (PCG noise per pixel) */
    unsigned int seed;
} bh_tone_params;

/* Radio tone map of px (W*H*4 floats, host memory) to rgb8 (W*H*3 bytes, row-major from the top
   left): colormap of (exposure * displayed intensity)^gamma, highlights burned to white, optional
   order tint, vignette, grain and dithered 8-bit rounding (k_tonemap, render.cu:360-407). */
int bh_tonemap(bh_ctx* ctx, const float* px, int W, int H, const bh_tone_params* params, unsigned char* rgb8);

/* HUD elements (HudGlyph, HudRect, renderer.hpp:61-62): a glyph of the font at (x, y) (top left of
   its cell, pixels) scaled by `scale`, colour rgba = R | G << 8 | B << 16 | A << 24; a rectangle
   [x0, x1) x [y0, y1) blended with its colour's alpha, drawn in increasing `layer`. */
typedef struct bh_hud_glyph {
    float x, y, scale;
    int glyph;
    uint32_t rgba;
} bh_hud_glyph;
typedef struct bh_hud_rect {
    int x0, y0, x1, y1;
    uint32_t rgba;
    int layer;
} bh_hud_rect;

/* A bitmap font: `glyphs` cells of cell_w x cell_h 8-bit coverage values, glyph-major then row-major
   (the reference compiles one in: FONT_ATLAS, common/font_atlas.h). advance (pixels per character
   at scale 1) and baseline are for the host-side text layout. */
typedef struct bh_font {
    const unsigned char* atlas;
    int cell_w, cell_h, glyphs;
    float advance;
    int baseline;
} bh_font;

/* Renderer::tonemap with its HUD (render.cu:811-846): the radio tone map of bh_tonemap, then the
   rectangles stable-sorted by layer, then every glyph's shadow (offset max(1.5, H / 720) px, black,
   3/4 of the glyph's alpha), then every glyph, bilinearly sampled from the font (k_hud_rects,
   k_hud_glyphs, render.cu:409-446). At most 1024 rectangles and 4096 glyphs are drawn, as the
   reference does. Each element is blended in list order (the reference's blocks blend overlapping
   elements in whatever order they run). Errors: font NULL with glyphs, a glyph index outside the
   font, a scale outside (0, 64]. */
int bh_tonemap_hud(bh_ctx* ctx, const float* px, int W, int H, const bh_tone_params* params,
                   const bh_hud_glyph* glyphs, int n_glyphs, const bh_hud_rect* rects, int n_rects,
                   const bh_font* font, unsigned char* rgb8);

/* Diagnostic: evaluates one elementary function of the device math library on n inputs (no context
   needed; device 0). fn: 0 exp, 1 log, 2 expm1, 3 cbrt, 4 x^-0.2, 5 sin, 6 cos, 7 atan2(x, y),
   8 acos, 9 sqrt, 10 x / y, 11 asin. Used by the tests that pin the device functions bit for bit. */
int bh_math_eval(int fn, const double* x, const double* y, double* out, int n);

/* Renderer::~Renderer                                           render.cu:526-529 */
void bh_destroy(bh_ctx* ctx);

/* Last error text of a context (or of the last failed bh_create when ctx is NULL). */
const char* bh_last_error(const bh_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
