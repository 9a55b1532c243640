/* turb_sheared.h - the reference's flow turbulence on gfx950 (device code, included by
   trace_fp64.h after TraceParams and pcg).

   Model (blackholes/src/emission.cuh:30-51, device_util.cuh:31-71): Perlin gradient noise with
   quintic fade and 12 cube-edge gradients picked by a nested PCG hash, 4-octave fBm (rotation
   (0.8, 0.6) in x-z, lacunarity 2.03, gain 0.55, scaled by 2.2 / sqrt(sum of the squared octave
   weights)), sheared by the local
   Keplerian rotation: each noise generation is born every tcorr orbital periods, advected with the
   flow, and cross-faded with the next. The density factor is exp(a n - a^2 / 2).

   Evaluation (render.cu:447-460,594-619,679-682, emission.cuh:112-123): for each frame the field
   is baked into nt slices of retarded time (slice k at t_ref(r) + d0 + k dd, t_ref(r) = t_cam -
   (r_cam - r + 2 ln(r_cam / r)): the time at which a radial light ray from r reaches the camera's
   distance), each a (z/R, ln R, phi) grid, phi fastest. A flow sample reads two slices with
   trilinear filtering and interpolates between them in its own emission time (slow light).

   Choices here: the slices are float arrays in HBM (453 MB at the reference's
   512 x 192 x 48 x 24; MI355X has 288 GB) filtered in software — wrap in phi, clamp in ln R and
   z/R, texel centres at (i + 0.5) / n, weights rounded to 1/256 like the 8-bit fraction of a
   texture unit's filter — so the oracle can restate every value; double-precision sine, cosine, log
   and exp through bh_math.h; 1 / sqrtf where the reference uses the approximate rsqrtf, and the
   bh_math exp where it uses __expf. The bake is one work-item per cell (1-D grid, phi fastest, so
   stores coalesce). */
#pragma once

// Parameters of one bake: the scene's noise constants and the frame's camera (render.cu:615-619)
struct ShearedBake {
    double a, r_isco;         // Keplerian rotation of the noise (emission.cuh:20)
    double cam_t, cam_r;      // retarded-time reference
    float lr0, lr1, zmax;     // ln R range, |z/R| range of the slices
    float d0, dd;             // retarded-time offset of slice 0 and slice spacing
    float amp, scale, tcorr;  // flow_turb_amp, flow_turb_scale, flow_tcorr
    uint32_t seed;            // scene seed ^ 0x51f15eed
    int nphi, nlr, nz, nt;
};

__device__ __forceinline__ float sh_lerp(float a, float b, float t) { return a + (b - a) * t; }

// gradient of lattice point (ix, iy, iz) dotted with the offset (fx, fy, fz) (device_util.cuh:31-40)
__device__ __forceinline__ float sh_grad(int ix, int iy, int iz, uint32_t seed, float fx, float fy, float fz) {
    const uint32_t h = pcg((uint32_t)ix * 73856093u ^ pcg((uint32_t)iy * 19349663u ^ pcg((uint32_t)iz * 83492791u ^ seed)));
    const uint32_t e = h % 12u;
    // edges of the cube: the two non-zero components and their signs
    const float p = e < 4u ? fx : (e < 8u ? fx : fy);
    const float q = e < 4u ? fy : fz;
    const float sp = (e & 1u) ? -p : p;
    const float sq = (e & 2u) ? -q : q;
    return sp + sq;
}

__device__ __forceinline__ float sh_fade(float u) { return u * u * u * (u * (u * 6.f - 15.f) + 10.f); }

// gradient noise, ~[-1, 1] (device_util.cuh:41-59)
__device__ __forceinline__ float sh_gnoise(float x, float y, float z, uint32_t seed) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const float ux = x - fx, uy = y - fy, uz = z - fz;
    const float wx = sh_fade(ux), wy = sh_fade(uy), wz = sh_fade(uz);
    const float n000 = sh_grad(ix, iy, iz, seed, ux, uy, uz);
    const float n100 = sh_grad(ix + 1, iy, iz, seed, ux - 1, uy, uz);
    const float n010 = sh_grad(ix, iy + 1, iz, seed, ux, uy - 1, uz);
    const float n110 = sh_grad(ix + 1, iy + 1, iz, seed, ux - 1, uy - 1, uz);
    const float n001 = sh_grad(ix, iy, iz + 1, seed, ux, uy, uz - 1);
    const float n101 = sh_grad(ix + 1, iy, iz + 1, seed, ux - 1, uy, uz - 1);
    const float n011 = sh_grad(ix, iy + 1, iz + 1, seed, ux, uy - 1, uz - 1);
    const float n111 = sh_grad(ix + 1, iy + 1, iz + 1, seed, ux - 1, uy - 1, uz - 1);
    const float x00 = sh_lerp(n000, n100, wx), x10 = sh_lerp(n010, n110, wx);
    const float x01 = sh_lerp(n001, n101, wx), x11 = sh_lerp(n011, n111, wx);
    return sh_lerp(sh_lerp(x00, x10, wy), sh_lerp(x01, x11, wy), wz);
}

// 4-octave fBm, 2.2 / sqrt(sum of squared weights) times the weighted octave sum (device_util.cuh:61-71)
__device__ __forceinline__ float sh_fbm(float x, float y, float z, uint32_t seed) {
    float s = 0.f, amp = 1.f, norm = 0.f;
    for (int i = 0; i < 4; ++i) {
        s += amp * sh_gnoise(x, y, z, seed + 1013u * (uint32_t)i);
        norm += amp * amp;
        const float nx = 0.80f * x + 0.60f * z, nz = -0.60f * x + 0.80f * z;
        x = nx * 2.03f;
        y = y * 2.03f;
        z = nz * 2.03f;
        amp *= 0.55f;
    }
    return s * 2.2f / __builtin_sqrtf(norm);
}

// sheared, finite-lifetime noise at cylindrical (R, phi, z) and time t (emission.cuh:34-51)
__device__ __forceinline__ float sh_noise(const ShearedBake& B, double R, double phi, double z, double t) {
    const double Rk = fmax(R, B.r_isco);
    const double Om = 1.0 / (Rk * sqrt(Rk) + B.a);
    const double P = 6.283185307179586 / Om * B.tcorr;   // decorrelation time
    const double s = t / P;
    const double k = floor(s);
    const float f = (float)(s - k);
    float acc = 0.f;
    for (int j = 0; j < 2; ++j) {
        const double tb = (k + j) * P;           // birth of this generation
        const double ph = phi - Om * (t - tb);   // comoving azimuth at birth
        const float inv = 1.f / (B.scale * (float)R);
        double sp, cp;
        bhm::sincos(ph, &sp, &cp);
        const float x = (float)(R * cp) * inv, y = (float)(R * sp) * inv, zz = (float)z * inv * 2.f;
        const float n = sh_fbm(x, y, zz, B.seed + 7919u * (uint32_t)(long long)(k + j));
        const float w = j == 0 ? 1.f - f : f;
        acc += w * n;
    }
    return acc * (1.f / __builtin_sqrtf((1.f - f) * (1.f - f) + f * f));
}

// Density factor of cell (i phi, j ln R, l z/R) of slice k (render.cu:449-460)
__device__ __forceinline__ float sh_cell(const ShearedBake& B, int i, int j, int l, int k) {
    const double phi = (i + 0.5) / B.nphi * 2.0 * PI;
    const double R = bhm::exp(B.lr0 + (j + 0.5) / B.nlr * (B.lr1 - B.lr0));
    const double z = (-1.0 + 2.0 * (l + 0.5) / B.nz) * B.zmax * R;
    const double r = sqrt(R * R + z * z);
    const float delta = B.d0 + (float)k * B.dd;
    const double t = B.cam_t - (B.cam_r - r + 2.0 * bhm::log(B.cam_r / r)) + delta;
    const float n = sh_noise(B, R, phi, z, t);
    const float a = B.amp;
    return (float)bhm::exp((double)(a * n - 0.5f * a * a));
}

// Trilinear read of one slice at normalised coordinates (x wraps; y, z clamp), texel centres at
// (i + 0.5) / n, weights rounded to 1/256.
__device__ __forceinline__ float sh_tex(const float* __restrict__ T, int nx, int ny, int nz, float x, float y, float z) {
    if (!(x == x)) x = 0.f;
    if (!(y == y)) y = 0.f;
    if (!(z == z)) z = 0.f;
    x = x - floorf(x);
    if (!(x >= 0.f && x <= 1.f)) x = 0.f;   // +-inf
    const float xb = x * (float)nx - 0.5f;
    const float yb = fminf(fmaxf(y * (float)ny - 0.5f, -1.f), (float)ny);
    const float zb = fminf(fmaxf(z * (float)nz - 0.5f, -1.f), (float)nz);
    const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
    const float ax = __builtin_rintf((xb - fx) * 256.f) * (1.f / 256.f);
    const float ay = __builtin_rintf((yb - fy) * 256.f) * (1.f / 256.f);
    const float az = __builtin_rintf((zb - fz) * 256.f) * (1.f / 256.f);
    const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    const int x0 = ix < 0 ? ix + nx : (ix >= nx ? ix - nx : ix);
    const int x1 = x0 + 1 == nx ? 0 : x0 + 1;
    const int y0 = min(max(iy, 0), ny - 1), y1 = min(max(iy + 1, 0), ny - 1);
    const int z0 = min(max(iz, 0), nz - 1), z1 = min(max(iz + 1, 0), nz - 1);
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const float v000 = T[z0 * sz + y0 * sy + x0], v100 = T[z0 * sz + y0 * sy + x1];
    const float v010 = T[z0 * sz + y1 * sy + x0], v110 = T[z0 * sz + y1 * sy + x1];
    const float v001 = T[z1 * sz + y0 * sy + x0], v101 = T[z1 * sz + y0 * sy + x1];
    const float v011 = T[z1 * sz + y1 * sy + x0], v111 = T[z1 * sz + y1 * sy + x1];
    const float bx = 1.f - ax, by = 1.f - ay, bz = 1.f - az;
    return bz * (by * (bx * v000 + ax * v100) + ay * (bx * v010 + ax * v110)) +
           az * (by * (bx * v001 + ax * v101) + ay * (bx * v011 + ax * v111));
}

// Density factor of a flow sample at (r, R = r sin theta, z = r cos theta, phi) and emission time t
// (emission.cuh:113-122), in the FP64 tracer's arithmetic (float slice coordinates as the reference).
__device__ __forceinline__ float sheared_factor(const TraceParams& P, double r, double R, double z, double phi,
                                                double t) {
    const ShearedView& T = *P.ts;
    const float u = (float)((bhm::log(R) - T.lr0) / (T.lr1 - T.lr0));
    const float v = (float)(phi * (0.5 / PI));
    const float w = (float)(0.5 + 0.5 * (z / R) / T.zmax);
    const double tref = T.cam_t - (T.cam_r - r + 2.0 * bhm::log(T.cam_r / r));
    float s = (float)((t - tref) - T.d0) / T.dd;
    s = fminf(fmaxf(s, 0.f), (float)(T.nt - 1) - 1e-4f);
    const int k = (int)s;
    const float f = s - (float)k;
    const size_t slice = (size_t)T.nphi * T.nlr * T.nz;
    const float A = sh_tex(T.data + (size_t)k * slice, T.nphi, T.nlr, T.nz, v, u, w);
    const float B = sh_tex(T.data + (size_t)(k + 1) * slice, T.nphi, T.nlr, T.nz, v, u, w);
    return A + (B - A) * f;
}
