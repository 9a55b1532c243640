/* sheared.cpp - CPU restatement of bhrt's flow turbulence: sheared, finite-lifetime fBm noise baked
   into retarded-time slices and read with trilinear filtering. TEST INFRASTRUCTURE ONLY (the same
   rules as oracle.h: only tests/ load it, as the checker of the GPU bake and of the tracer's
   lookup).

   Follows, operation by operation in the same types (float where the reference computes in float,
   double where it does), with -ffp-contract=off:
   - pcg hash, gradient noise with quintic fade and 12 cube-edge gradients, 4-octave fBm:
     blackholes/src/device_util.cuh:15-19,31-71;
   - sheared noise (generations born every tcorr orbital periods, advected with the Keplerian
     rotation, cross-faded): emission.cuh:20,34-51; retarded reference time: emission.cuh:26-28;
   - the bake of one cell: render.cu:449-460, slice parameters render.cu:615-619;
   - the lookup: emission.cuh:113-122 with the texture filter restated in software (wrap in phi,
     clamp in ln R and z/R, texel centres at (i + 0.5) / n, weights rounded to 1/256).
   Deliberate differences from the reference, shared with the GPU code: sine, cosine, exp and log in
   double through this oracle's elementary functions, 1 / sqrtf for rsqrtf, the double exp for
   __expf, and the texture filter above (hardware filtering cannot be restated exactly).
   Parity with the reference's own values is unpinned (its noise is device code that cannot be
   compiled here); the tests pin what the mathematics fixes. */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "../spacetime-optics_amd/csrc/bh_math.h"

#if ORACLE_BHM
#define S_EXP bhm::exp
#define S_LOG bhm::log
static void s_sincos(double x, double* s, double* c) { bhm::sincos(x, s, c); }
#else
#define S_EXP std::exp
#define S_LOG std::log
static void s_sincos(double x, double* s, double* c) {
    *s = std::sin(x);
    *c = std::cos(x);
}
#endif

namespace {

const double PI_D = 3.14159265358979323846;

uint32_t hash(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}

float lerpf(float a, float b, float t) { return a + (b - a) * t; }

float grad(int ix, int iy, int iz, uint32_t seed, float fx, float fy, float fz) {
    uint32_t h = hash((uint32_t)ix * 73856093u ^ hash((uint32_t)iy * 19349663u ^ hash((uint32_t)iz * 83492791u ^ seed)));
    switch (h % 12u) {
        case 0: return fx + fy;
        case 1: return -fx + fy;
        case 2: return fx - fy;
        case 3: return -fx - fy;
        case 4: return fx + fz;
        case 5: return -fx + fz;
        case 6: return fx - fz;
        case 7: return -fx - fz;
        case 8: return fy + fz;
        case 9: return -fy + fz;
        case 10: return fy - fz;
        default: return -fy - fz;
    }
}

float fade(float u) { return u * u * u * (u * (u * 6.f - 15.f) + 10.f); }

float gnoise(float x, float y, float z, uint32_t seed) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float ux = x - fx, uy = y - fy, uz = z - fz;
    float wx = fade(ux), wy = fade(uy), wz = fade(uz);
    float c[2][2][2];
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i)
                c[k][j][i] = grad(ix + i, iy + j, iz + k, seed, i ? ux - 1 : ux, j ? uy - 1 : uy, k ? uz - 1 : uz);
    float x00 = lerpf(c[0][0][0], c[0][0][1], wx), x10 = lerpf(c[0][1][0], c[0][1][1], wx);
    float x01 = lerpf(c[1][0][0], c[1][0][1], wx), x11 = lerpf(c[1][1][0], c[1][1][1], wx);
    return lerpf(lerpf(x00, x10, wy), lerpf(x01, x11, wy), wz);
}

float fbm(float x, float y, float z, int oct, uint32_t seed) {
    float s = 0.f, amp = 1.f, norm = 0.f;
    for (int i = 0; i < oct; ++i) {
        s += amp * gnoise(x, y, z, seed + 1013u * (uint32_t)i);
        norm += amp * amp;
        float nx = 0.80f * x + 0.60f * z, nz = -0.60f * x + 0.80f * z;
        x = nx * 2.03f;
        y = y * 2.03f;
        z = nz * 2.03f;
        amp *= 0.55f;
    }
    return s * 2.2f / std::sqrt(norm);
}

float noise(const or_sheared* B, double R, double phi, double z, double t) {
    double Rk = std::fmax(R, B->r_isco);
    double Om = 1.0 / (Rk * std::sqrt(Rk) + B->a);
    double P = 6.283185307179586 / Om * B->tcorr;
    double s = t / P;
    double k = std::floor(s);
    float f = (float)(s - k);
    float acc = 0.f;
    for (int j = 0; j < 2; ++j) {
        double tb = (k + j) * P;
        double ph = phi - Om * (t - tb);
        float inv = 1.f / (B->scale * (float)R);
        double sp, cp;
        s_sincos(ph, &sp, &cp);
        float x = (float)(R * cp) * inv, y = (float)(R * sp) * inv, zz = (float)z * inv * 2.f;
        float n = fbm(x, y, zz, 4, B->seed + 7919u * (uint32_t)(long long)(k + j));
        float w = j == 0 ? 1.f - f : f;
        acc += w * n;
    }
    return acc * (1.f / std::sqrt((1.f - f) * (1.f - f) + f * f));
}

float tex(const float* T, int nx, int ny, int nz, float x, float y, float z) {
    if (std::isnan(x)) x = 0.f;
    if (std::isnan(y)) y = 0.f;
    if (std::isnan(z)) z = 0.f;
    x = x - std::floor(x);
    if (!(x >= 0.f && x <= 1.f)) x = 0.f;
    float xb = x * (float)nx - 0.5f;
    float yb = std::fmin(std::fmax(y * (float)ny - 0.5f, -1.f), (float)ny);
    float zb = std::fmin(std::fmax(z * (float)nz - 0.5f, -1.f), (float)nz);
    float fx = std::floor(xb), fy = std::floor(yb), fz = std::floor(zb);
    float a[3] = {std::nearbyint((xb - fx) * 256.f) * (1.f / 256.f), std::nearbyint((yb - fy) * 256.f) * (1.f / 256.f),
                  std::nearbyint((zb - fz) * 256.f) * (1.f / 256.f)};
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    int xs[2], ys[2], zs[2];
    xs[0] = ((ix % nx) + nx) % nx;
    xs[1] = (xs[0] + 1) % nx;
    ys[0] = std::min(std::max(iy, 0), ny - 1);
    ys[1] = std::min(std::max(iy + 1, 0), ny - 1);
    zs[0] = std::min(std::max(iz, 0), nz - 1);
    zs[1] = std::min(std::max(iz + 1, 0), nz - 1);
    auto at = [&](int i, int j, int k) { return T[((size_t)zs[k] * ny + ys[j]) * nx + xs[i]]; };
    float b[3] = {1.f - a[0], 1.f - a[1], 1.f - a[2]};
    return b[2] * (b[1] * (b[0] * at(0, 0, 0) + a[0] * at(1, 0, 0)) + a[1] * (b[0] * at(0, 1, 0) + a[0] * at(1, 1, 0))) +
           a[2] * (b[1] * (b[0] * at(0, 0, 1) + a[0] * at(1, 0, 1)) + a[1] * (b[0] * at(0, 1, 1) + a[0] * at(1, 1, 1)));
}

}  // namespace

extern "C" {

void or_sheared_params(const or_scene* S, double cam_t, double cam_r, const int dims[4], or_sheared* B) {
    B->a = S->a;
    B->r_isco = S->r_isco;
    B->cam_t = cam_t;
    B->cam_r = cam_r;
    B->lr0 = (float)std::log(0.8 * S->rh);
    B->lr1 = (float)std::log(S->r_flow_max * 1.05);
    B->zmax = (float)std::min(2.0, 5.0 * S->h_r);
    const int def[4] = {512, 192, 48, 24};
    B->nphi = dims && dims[0] > 0 ? dims[0] : def[0];
    B->nlr = dims && dims[1] > 0 ? dims[1] : def[1];
    B->nz = dims && dims[2] > 0 ? dims[2] : def[2];
    B->nt = dims && dims[3] > 0 ? dims[3] : def[3];
    B->d0 = -250.f;
    B->dd = 260.f / (float)(B->nt - 1);
    B->amp = S->flow_turb_amp;
    B->scale = S->flow_turb_scale;
    B->tcorr = S->flow_tcorr;
    B->seed = S->seed ^ 0x51f15eedu;
}

float or_sheared_gnoise(float x, float y, float z, uint32_t seed) { return gnoise(x, y, z, seed); }
float or_sheared_fbm(float x, float y, float z, int octaves, uint32_t seed) { return fbm(x, y, z, octaves, seed); }
float or_sheared_noise(const or_sheared* B, double R, double phi, double z, double t) { return noise(B, R, phi, z, t); }
float or_sheared_tex(const float* T, int nx, int ny, int nz, float x, float y, float z) {
    return tex(T, nx, ny, nz, x, y, z);
}

float or_sheared_cell(const or_sheared* B, int i, int j, int l, int k) {
    double phi = (i + 0.5) / B->nphi * 2.0 * PI_D;
    double R = S_EXP(B->lr0 + (j + 0.5) / B->nlr * (B->lr1 - B->lr0));
    double z = (-1.0 + 2.0 * (l + 0.5) / B->nz) * B->zmax * R;
    double r = std::sqrt(R * R + z * z);
    float delta = B->d0 + (float)k * B->dd;
    double t = B->cam_t - (B->cam_r - r + 2.0 * S_LOG(B->cam_r / r)) + delta;
    float n = noise(B, R, phi, z, t);
    float a = B->amp;
    return (float)S_EXP((double)(a * n - 0.5f * a * a));
}

void or_sheared_bake(const or_sheared* B, float* out, int nthreads) {
    const long long total = (long long)B->nphi * B->nlr * B->nz * B->nt;
    const long long rows = total / B->nphi;
    std::atomic<long long> next(0);
    auto work = [&]() {
        for (;;) {
            long long row = next.fetch_add(1);
            if (row >= rows) break;
            long long rest = row;
            int j = (int)(rest % B->nlr);
            rest /= B->nlr;
            int l = (int)(rest % B->nz);
            int k = (int)(rest / B->nz);
            for (int i = 0; i < B->nphi; ++i) out[row * B->nphi + i] = or_sheared_cell(B, i, j, l, k);
        }
    };
    int n = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> th;
    for (int t = 0; t < n; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
}

// The tracer's lookup (emission.cuh:113-122): density factor at (r, R, z, phi) and emission time t.
float or_sheared_factor(const or_scene* S, double r, double R, double z, double phi, double t) {
    const or_sheared* B = &S->ts;
    float u = (float)((S_LOG(R) - B->lr0) / (B->lr1 - B->lr0));
    float v = (float)(phi * (0.5 / PI_D));
    float w = (float)(0.5 + 0.5 * (z / R) / B->zmax);
    double tref = B->cam_t - (B->cam_r - r + 2.0 * S_LOG(B->cam_r / r));
    float s = (float)((t - tref) - B->d0) / B->dd;
    s = std::fmin(std::fmax(s, 0.f), (float)(B->nt - 1) - 1e-4f);
    int k = (int)s;
    float f = s - (float)k;
    size_t slice = (size_t)B->nphi * B->nlr * B->nz;
    float A = tex(S->ts_slices + (size_t)k * slice, B->nphi, B->nlr, B->nz, v, u, w);
    float Bv = tex(S->ts_slices + (size_t)(k + 1) * slice, B->nphi, B->nlr, B->nz, v, u, w);
    return A + (Bv - A) * f;
}

}  // extern "C"
