/* scene.h - shot .ini -> bh_scene + tables (blackholes/src/scene_build.cpp:83-214 of the reference,
   integrator / units / flow / K2 parts; disk, sky and colour are out of scope). */
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/bhrt.h"
#include "config.h"

namespace bh {

// blackholes/src/common.hpp:17-29
namespace cgs {
constexpr double C = 2.99792458e10;
constexpr double G = 6.67430e-8;
constexpr double MSUN = 1.98847e33;
constexpr double PC = 3.0856775814913673e18;
constexpr double JY = 1e-23;
}  // namespace cgs
constexpr double PI = 3.14159265358979323846;
constexpr double RAD2UAS = 180.0 / PI * 3600.0e6;

// SceneInfo (scene_build.hpp:8-18)
struct SceneInfo {
    double mass_msun = 0, distance_cm = 0, rg_cm = 0, tg_s = 0, uas_per_m = 0;
    double radio_flux_target_jy = 0;
    double normalize_fov_m = 40;
};

// SceneTables (renderer.hpp:19-23) + the turbulence grid (new input; see bhrt.h bh_tables)
struct SceneTables {
    std::vector<float> k2_ln;
    float k2_l0 = 0, k2_dl = 0;
    std::vector<float> turb;
    int turb_n = 0, turb_nt = 0;
    double turb_half = 0, turb_t0 = 0, turb_dt = 0;
    int turb_model = BH_TURB_GRID;
    int ts_dims[4] = {0, 0, 0, 0};
    bh_tables view() const;   // non-owning C view
};

// Tracer choices from the [tracer] section (new keys): precision = fp64 | mixed,
// emissivity = exact | mlp, jitter = 0 | 1.
struct TracerOptions {
    int precision = BH_PREC_FP64;
    int emissivity = BH_EMIS_EXACT;
    int jitter = 1;
};

double horizon_radius(double a);   // kerr.hpp:44
double isco_radius(double a);      // kerr.hpp:46-50

// ln K2(1/Theta) on n points of ln Theta in [ln 0.02, ln 1e4] (scene_build.cpp:159-170)
void k2_table(int n, std::vector<float>* tab, float* l0, float* dl);

/* Seeded log-normal density factor exp(sigma g - sigma^2/2) on nt time slices of a cell-centred
   n^3 grid over [-half, half]^3 (x fastest). g is unit-variance Gaussian noise with correlation
   exp(-d^2 / (2 corr_m^2)) in space and exp(-dt^2 / (2 corr_t^2)) between slices (periodic
   boundaries), made by separable Gaussian smoothing of counter-based white noise. */
void make_turbulence(int n, int nt, double half, double sigma, double corr_m, double corr_t, double dt,
                     uint32_t seed, std::vector<float>* out);

// with_grid = false skips building the turbulence grid (scalars and the K2 table only)
bool build_scene(const Config& c, bh_scene* S, SceneTables* T, SceneInfo* info, TracerOptions* opt,
                 std::string* err, bool with_grid = true);

}  // namespace bh
