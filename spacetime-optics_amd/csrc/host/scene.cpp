#include "scene.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <sstream>
#include <thread>

namespace bh {

bh_tables SceneTables::view() const {
    bh_tables t;
    std::memset(&t, 0, sizeof t);
    t.k2_ln = k2_ln.empty() ? nullptr : k2_ln.data();
    t.k2_n = (int)k2_ln.size();
    t.k2_l0 = k2_l0;
    t.k2_dl = k2_dl;
    t.turb = turb.empty() ? nullptr : turb.data();
    t.turb_n = turb_n;
    t.turb_nt = turb_nt;
    t.turb_half = turb_half;
    t.turb_t0 = turb_t0;
    t.turb_dt = turb_dt;
    t.turb_model = turb_model;
    for (int i = 0; i < 4; ++i) t.ts_dims[i] = ts_dims[i];
    return t;
}

double horizon_radius(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

double isco_radius(double a) {
    double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
    double z2 = std::sqrt(3.0 * a * a + z1 * z1);
    return 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

void k2_table(int n, std::vector<float>* tab, float* l0f, float* dlf) {
    const double l0 = std::log(0.02), l1 = std::log(1e4);
    tab->resize(n);
    for (int i = 0; i < n; ++i) {
        double th = std::exp(l0 + (l1 - l0) * i / (n - 1));
        (*tab)[i] = (float)std::log(std::cyl_bessel_k(2.0, 1.0 / th));
    }
    *l0f = (float)l0;
    *dlf = (float)((l1 - l0) / (n - 1));
}

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Unit normal deviate number i of stream `seed` (Box-Muller on two counter-based uniforms)
double normal_deviate(uint64_t seed, uint64_t i) {
    uint64_t k = splitmix64(seed * 0x2545F4914F6CDD1Dull + 0x1234567ull);
    double u1 = ((splitmix64(k ^ (2 * i)) >> 11) + 0.5) * 0x1p-53;
    double u2 = (splitmix64(k ^ (2 * i + 1)) >> 11) * 0x1p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * PI * u2);
}

std::vector<double> gauss_kernel(double sigma_cells, int* radius) {
    int r = std::max(1, (int)std::ceil(4.0 * sigma_cells));
    std::vector<double> g(2 * r + 1);
    double s = 0;
    for (int k = -r; k <= r; ++k) s += (g[k + r] = std::exp(-0.5 * k * k / (sigma_cells * sigma_cells)));
    for (double& v : g) v /= s;
    *radius = r;
    return g;
}

// periodic 1D convolution of `len` lines along a stride; lines are enumerated by (outer, inner)
void smooth_axis(std::vector<double>& f, size_t count, size_t len, size_t stride, const std::vector<double>& g,
                 int r, int nthreads) {
    // line l: base = (l / stride) * stride * len + (l % stride)
    auto work = [&](size_t l0, size_t l1) {
        std::vector<double> line(len), out(len);
        for (size_t l = l0; l < l1; ++l) {
            size_t base = (l / stride) * stride * len + (l % stride);
            for (size_t i = 0; i < len; ++i) line[i] = f[base + i * stride];
            for (size_t i = 0; i < len; ++i) {
                double acc = 0;
                for (int k = -r; k <= r; ++k) {
                    long j = ((long)i + k) % (long)len;
                    if (j < 0) j += (long)len;
                    acc += g[k + r] * line[(size_t)j];
                }
                out[i] = acc;
            }
            for (size_t i = 0; i < len; ++i) f[base + i * stride] = out[i];
        }
    };
    std::vector<std::thread> th;
    size_t per = (count + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        size_t a = t * per, b = std::min(count, a + per);
        if (a < b) th.emplace_back(work, a, b);
    }
    for (auto& t : th) t.join();
}

// Black-hole frame -> galactic rotation of the celestial sphere (scene_build.cpp:181-204). Two
// orthonormal bases are matched: the camera's (the direction straight behind the hole, the sky's up
// seen from there, and their cross product) and the galactic one around the direction (l, b) (that
// direction, its north turned by roll, and their cross product); M = G B^T maps the first onto the
// second. Same operation order as the reference (its shot fixtures pin it).
void sky_rotation(double incl, double phi, double l, double b, double roll, double M[9]) {
    using A3 = std::array<double, 3>;
    auto dot3 = [](const A3& p, const A3& q) { return p[0] * q[0] + p[1] * q[1] + p[2] * q[2]; };
    auto minus_along = [](const A3& p, const A3& v, double k) { return A3{p[0] - v[0] * k, p[1] - v[1] * k, p[2] - v[2] * k}; };
    auto unit = [&](const A3& v) {
        const double k = 1.0 / std::sqrt(dot3(v, v));
        return A3{v[0] * k, v[1] * k, v[2] * k};
    };
    auto cross3 = [](const A3& p, const A3& q) {
        return A3{p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]};
    };
    const A3 pole = {0.0, 0.0, 1.0};
    const A3 behind = {-(std::sin(incl) * std::cos(phi)), -(std::sin(incl) * std::sin(phi)), -std::cos(incl)};
    A3 up = unit(minus_along(pole, behind, dot3(pole, behind)));
    if (!(std::sqrt(dot3(up, up)) > 0.5)) up = {1.0, 0.0, 0.0};   // camera on the pole
    const A3 side = cross3(up, behind);
    const A3 gt = {std::cos(b) * std::cos(l), std::cos(b) * std::sin(l), std::sin(b)};
    const A3 gn = unit(minus_along(pole, gt, dot3(pole, gt)));
    const A3 gs = cross3(gn, gt);
    const double cr = std::cos(roll), sr = std::sin(roll);
    const A3 gu = {gn[0] * cr + gs[0] * sr, gn[1] * cr + gs[1] * sr, gn[2] * cr + gs[2] * sr};
    const A3 gv = cross3(gu, gt);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[3 * i + j] = gt[i] * behind[j] + gu[i] * up[j] + gv[i] * side[j];
}

}  // namespace

void make_turbulence(int n, int nt, double half, double sigma, double corr_m, double corr_t, double dt,
                     uint32_t seed, std::vector<float>* out) {
    nt = std::max(nt, 1);
    const size_t N = (size_t)n, vol = N * N * N, total = vol * nt;
    std::vector<double> f(total);
    for (size_t i = 0; i < total; ++i) f[i] = normal_deviate(seed, i);
    int nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const double dx = 2.0 * half / n;
    // correlation exp(-d^2/(2 l^2)) of smoothed white noise needs a kernel of std l / sqrt(2)
    int r;
    std::vector<double> g = gauss_kernel(corr_m / std::sqrt(2.0) / dx, &r);
    double g2 = 0;
    for (double v : g) g2 += v * v;
    double var = g2 * g2 * g2;
    smooth_axis(f, total / N, N, 1, g, r, nthreads);            // x
    smooth_axis(f, total / N, N, N, g, r, nthreads);            // y
    smooth_axis(f, total / N, N, N * N, g, r, nthreads);        // z
    if (nt > 1 && corr_t > 0 && dt > 0) {
        // the periodic time axis must hold the whole kernel (radius 4 sigma)
        double st = std::min(corr_t / std::sqrt(2.0) / dt, (nt - 1) / 8.0);
        int rt = 0;
        std::vector<double> gt = st > 0.25 ? gauss_kernel(st, &rt) : std::vector<double>{1.0};
        double gt2 = 0;
        for (double v : gt) gt2 += v * v;
        var *= gt2;
        smooth_axis(f, vol, (size_t)nt, vol, gt, rt, nthreads);   // t
    }
    const double inv = 1.0 / std::sqrt(var);
    out->resize(total);
    for (size_t i = 0; i < total; ++i) (*out)[i] = (float)std::exp(sigma * (f[i] * inv) - 0.5 * sigma * sigma);
}

// blackholes/src/scene_build.cpp:83-105 (integrator, units) and 143-170 (flow, K2)
bool build_scene(const Config& c, bh_scene* S, SceneTables* T, SceneInfo* info, TracerOptions* opt,
                 std::string* err, bool with_grid) {
    std::memset(S, 0, sizeof(*S));
    double a = c.num("bh.spin", 0.94);
    if (a <= 0.0) a = 1e-6;
    if (a > 0.9999) a = 0.9999;
    S->a = a;
    S->rh = horizon_radius(a);
    S->r_isco = isco_radius(a);
    S->tol = c.num("render.tolerance", 1e-9);
    S->max_steps = c.integer("render.max_steps", 20000);
    S->r_escape = c.num("render.r_escape", 2e4);
    S->r_stop = S->rh + c.num("render.horizon_margin", 0.01);
    S->step_frac = c.num("render.step_fraction", 0.04);
    std::string mode = c.str("render.mode", "visible");
    S->mode = mode == "radio" ? BH_MODE_RADIO : BH_MODE_VISIBLE;
    S->seed = (uint32_t)c.integer("render.seed", 1) * 2654435761u;
    if (!(S->tol > 0.0) || S->max_steps <= 0) {
        if (err) *err = "render.tolerance and render.max_steps must be positive";
        return false;
    }

    info->mass_msun = c.num("bh.mass_msun", 6.5e9);
    double dist_pc = c.has("bh.distance_mpc") ? c.num("bh.distance_mpc", 16.8) * 1e6 : c.num("bh.distance_pc", 16.8e6);
    info->distance_cm = dist_pc * cgs::PC;
    info->rg_cm = cgs::G * info->mass_msun * cgs::MSUN / (cgs::C * cgs::C);
    info->tg_s = info->rg_cm / cgs::C;
    info->uas_per_m = info->rg_cm / info->distance_cm * RAD2UAS;

    S->flow_on = c.flag("flow.enabled", S->mode == BH_MODE_RADIO) ? 1 : 0;
    S->n0 = c.num("flow.n0", 1e5);
    S->n_index = c.num("flow.n_index", 1.1);
    S->theta0 = c.num("flow.theta0", 25.0);
    S->theta_index = c.num("flow.theta_index", 0.84);
    S->h_r = c.num("flow.h_over_r", 0.4);
    S->beta_p = c.num("flow.beta", 10.0);
    S->r_flow_max = c.num("flow.r_max", 60.0);
    S->nu_obs = c.num("flow.nu_ghz", 230.0) * 1e9;
    S->rg_cm = info->rg_cm;
    S->flow_kep = c.num("flow.keplerian_fraction", 0.9);
    S->flow_turb_amp = (float)c.num("flow.turbulence", 0.7);
    S->flow_turb_scale = (float)c.num("flow.turbulence_scale", 0.45);
    S->flow_tcorr = (float)c.num("flow.correlation_orbits", 0.5);
    // Celestial sphere (visible mode): the reference's [sky] orientation keys (scene_build.cpp:181-204)
    // and brightness; a procedural checkerboard of sky.checker = nu, nv squares (default 24, 12:
    // 15 degree squares; 0 turns the sky off) in place of its star-map file
    {
        auto rad = [&](const char* key, double dflt) { return c.num(key, dflt) * PI / 180; };   // (deg * pi) / 180
        sky_rotation(rad("camera.incl", 80.0), rad("camera.phi", 0.0), rad("sky.l_deg", 0.0), rad("sky.b_deg", 0.0),
                     rad("sky.roll_deg", 0.0), S->sky_rot);
        S->sky_scale = (float)c.num("sky.brightness", 1.0);
        std::stringstream ss(c.str("sky.checker", "24, 12"));
        std::string item;
        int n[2] = {0, 0};
        for (int i = 0; i < 2 && std::getline(ss, item, ','); ++i) n[i] = std::atoi(item.c_str());
        if (n[1] <= 0) n[1] = n[0] / 2;
        if (n[0] < 0 || n[1] < 0 || n[0] > 4096 || n[1] > 4096) {
            if (err) *err = "sky.checker must be two counts in [0, 4096]";
            return false;
        }
        S->sky_nu = n[0];
        S->sky_nv = n[1];
        S->sky_on = S->mode == BH_MODE_VISIBLE && n[0] > 0 && n[1] > 0 ? 1 : 0;
    }
    info->radio_flux_target_jy = c.num("flow.flux_jy", 0.0);
    info->normalize_fov_m = c.num("flow.normalize_fov_m", 40.0);
    k2_table(512, &T->k2_ln, &T->k2_l0, &T->k2_dl);

    // Turbulence (new [turbulence] keys; amplitude from the reference key flow.turbulence):
    // model = sheared (default: the reference's baked sheared noise, baked by the library per frame;
    // cells = nphi,nlnR,nz and slices override its 512,192,48 x 24) or grid (a seeded log-normal
    // Cartesian grid made here)
    T->turb.clear();
    T->turb_n = T->turb_nt = 0;
    T->turb_model = BH_TURB_GRID;
    for (int i = 0; i < 4; ++i) T->ts_dims[i] = 0;
    const std::string model = c.str("turbulence.model", "sheared");
    if (model != "sheared" && model != "grid") {
        if (err) *err = "turbulence.model must be sheared or grid";
        return false;
    }
    if (model == "sheared") {
        T->turb_model = BH_TURB_SHEARED;
        std::stringstream ss(c.str("turbulence.cells", ""));
        std::string item;
        for (int i = 0; i < 3 && std::getline(ss, item, ','); ++i) T->ts_dims[i] = std::atoi(item.c_str());
        T->ts_dims[3] = c.integer("turbulence.slices", 0);
    } else if (with_grid && S->mode == BH_MODE_RADIO && S->flow_on && S->flow_turb_amp > 0.f) {
        int n = c.integer("turbulence.grid", 128);
        int nt = std::max(1, c.integer("turbulence.slices", 1));
        if (n < 2 || n > 1024) {
            if (err) *err = "turbulence.grid must be in [2, 1024]";
            return false;
        }
        T->turb_n = n;
        T->turb_nt = nt;
        T->turb_half = c.num("turbulence.half_m", S->r_flow_max);
        T->turb_t0 = c.num("turbulence.t0_m", 0.0);
        T->turb_dt = c.num("turbulence.dt_m", 10.0);
        make_turbulence(n, nt, T->turb_half, S->flow_turb_amp, c.num("turbulence.corr_m", 5.0),
                        c.num("turbulence.corr_t_m", 30.0), T->turb_dt, (uint32_t)c.integer("turbulence.seed", 1),
                        &T->turb);
    }
    if (opt) {
        std::string p = c.str("tracer.precision", "fp64");
        if (p != "fp64" && p != "mixed") {
            if (err) *err = "tracer.precision must be fp64 or mixed";
            return false;
        }
        opt->precision = p == "mixed" ? BH_PREC_MIXED : BH_PREC_FP64;
        std::string e = c.str("tracer.emissivity", "exact");
        if (e != "exact" && e != "mlp") {
            if (err) *err = "tracer.emissivity must be exact or mlp";
            return false;
        }
        opt->emissivity = e == "mlp" ? BH_EMIS_MLP : BH_EMIS_EXACT;
        opt->jitter = c.flag("tracer.jitter", true) ? 1 : 0;
    }
    return true;
}

}  // namespace bh
