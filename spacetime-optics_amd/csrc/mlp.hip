/* mlp.hip - thermal-synchrotron surrogate on gfx950: exact-routine sampler, BF16-MFMA training
   (forward + backward + Adam) and inference kernels, and the deployment table with its exact-routine
   judge. C ABI in include/bhrt.h (bh_mlp_*).

   Matrix products of the two 64 x 64 hidden layers use v_mfma_f32_16x16x16_bf16 with FP32
   accumulation. Each FP32 operand is split into a BF16 head and a BF16 tail (x = hi + lo, both
   rounded to nearest even) and a product is accumulated as lo*hi + hi*lo + hi*hi: three MFMAs that
   keep about 16 significant bits of every operand, so BF16 rounding does not bound the surrogate's
   accuracy. The 3 -> 64 input layer and the 64 -> 2 output layer run on the FP32 VALU.

   Work layout: one wave64 per workgroup; a wave processes tiles of 16 samples. MFMA operand layout
   (16x16x16, wave64): A[i][k] lane i + 16*(k/4), element k % 4; B[k][j] lane j + 16*(k/4), element
   k % 4; D[i][j] lane j + 16*(i/4), element i % 4. Activations go through LDS between layers to
   change from the D layout to the A layout. */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "trace_fp64.h"

using namespace bhmlp;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TM = 16;           // samples per tile
constexpr int LDA = NH + 4;      // padded row of an activation tile (floats)
constexpr int LDW = NH + 8;      // padded row of a BF16 weight matrix (halves)

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
typedef float fl2 __attribute__((ext_vector_type(2)));
typedef unsigned int u2 __attribute__((ext_vector_type(2)));

// two floats to packed BF16 (a in the low half), rounded to nearest even: one v_cvt_pk_bf16_f32
__device__ __forceinline__ unsigned int bf16x2(float a, float b) {
    return __builtin_bit_cast(unsigned int, __builtin_convertvector((fl2){a, b}, bf2));
}
// the low / high BF16 of a packed pair as a float
__device__ __forceinline__ float bf16_lo(unsigned int p) { return __uint_as_float(p << 16); }
__device__ __forceinline__ float bf16_hi(unsigned int p) { return __uint_as_float(p & 0xffff0000u); }

// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi), four values at once
__device__ __forceinline__ void split4(const float v[4], s4& hi, s4& lo) {
    const unsigned int h01 = bf16x2(v[0], v[1]), h23 = bf16x2(v[2], v[3]);
    const unsigned int l01 = bf16x2(v[0] - bf16_lo(h01), v[1] - bf16_hi(h01));
    const unsigned int l23 = bf16x2(v[2] - bf16_lo(h23), v[3] - bf16_hi(h23));
    hi = __builtin_bit_cast(s4, (u2){h01, h23});
    lo = __builtin_bit_cast(s4, (u2){l01, l23});
}

__device__ __forceinline__ f4 mfma3(s4 ah, s4 al, s4 bh, s4 bl, f4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
}

// logistic sigmoid from the hardware exp2 and reciprocal (about 1 ulp each; the library expf and
// the IEEE division cost ~30 instructions per element, on ~100 elements per lane and tile)
__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.f + z * (1.f - s));
}

struct alignas(16) WeightsLDS {   // 16-byte aligned: load_weights stores BF16 pairs as 32-bit words
    unsigned short w2h[NH * LDW], w2l[NH * LDW], w3h[NH * LDW], w3l[NH * LDW];
};
struct TileLDS {
    float x[TM][4];
    float z1[TM * LDA], h1[TM * LDA], z2[TM * LDA], h2[TM * LDA], z3[TM * LDA], h3[TM * LDA];
    float g0[TM * LDA], g1[TM * LDA];   // back-propagated gradients (ping-pong)
    float y[TM][2], t[TM][2], w[TM];
};

// split the FP32 master W2, W3 (row-major [k][j]) into BF16 head/tail arrays in LDS, two adjacent
// j at a time (LDW is even, so each pair is one aligned 32-bit store)
__device__ void load_weights(const float* __restrict__ P, WeightsLDS& W) {
    for (int e = 2 * threadIdx.x; e < NH * NH; e += 2 * blockDim.x) {
        const int k = e / NH, j = e % NH;
        const float a0 = P[OFF_W2 + e], a1 = P[OFF_W2 + e + 1], b0 = P[OFF_W3 + e], b1 = P[OFF_W3 + e + 1];
        const unsigned int ah = bf16x2(a0, a1), bh = bf16x2(b0, b1);
        *(unsigned int*)&W.w2h[k * LDW + j] = ah;
        *(unsigned int*)&W.w2l[k * LDW + j] = bf16x2(a0 - bf16_lo(ah), a1 - bf16_hi(ah));
        *(unsigned int*)&W.w3h[k * LDW + j] = bh;
        *(unsigned int*)&W.w3l[k * LDW + j] = bf16x2(b0 - bf16_lo(bh), b1 - bf16_hi(bh));
    }
}

// D(16 x 64) = A(16 x 64, LDS rows of LDA floats) * W (64 x 64 BF16 head/tail, [k][j])
__device__ __forceinline__ void gemm_fwd(const float* A, const unsigned short* Wh, const unsigned short* Wl, f4 acc[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int jb = 0; jb < 4; ++jb) acc[jb] = f4{0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < 4; ++kb) {
        const int k0 = 16 * kb + 4 * g;
        float av[4] = {A[r * LDA + k0], A[r * LDA + k0 + 1], A[r * LDA + k0 + 2], A[r * LDA + k0 + 3]};
        s4 ah, al;
        split4(av, ah, al);
        for (int jb = 0; jb < 4; ++jb) {
            const int j = 16 * jb + r;
            s4 bh, bl;
            for (int e = 0; e < 4; ++e) {
                bh[e] = (short)Wh[(k0 + e) * LDW + j];
                bl[e] = (short)Wl[(k0 + e) * LDW + j];
            }
            acc[jb] = mfma3(ah, al, bh, bl, acc[jb]);
        }
    }
}

// D(16 x 64)[i][k] = sum_j G[i][j] W[k][j]  (G: LDS rows of LDA floats; W: BF16 head/tail [k][j])
__device__ __forceinline__ void gemm_bwd(const float* G, const unsigned short* Wh, const unsigned short* Wl, f4 acc[4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int kb = 0; kb < 4; ++kb) acc[kb] = f4{0.f, 0.f, 0.f, 0.f};
    for (int jb = 0; jb < 4; ++jb) {
        const int j0 = 16 * jb + 4 * g;
        float av[4] = {G[r * LDA + j0], G[r * LDA + j0 + 1], G[r * LDA + j0 + 2], G[r * LDA + j0 + 3]};
        s4 ah, al;
        split4(av, ah, al);
        for (int kb = 0; kb < 4; ++kb) {
            const int k = 16 * kb + r;
            s4 bh, bl;
            for (int e = 0; e < 4; ++e) {
                bh[e] = (short)Wh[k * LDW + j0 + e];
                bl[e] = (short)Wl[k * LDW + j0 + e];
            }
            acc[kb] = mfma3(ah, al, bh, bl, acc[kb]);
        }
    }
}

// gW[m][j] += sum_i H[i][m] G[i][j] over the 16 samples of a tile; gW in D layout [mb][jb]
__device__ __forceinline__ void gemm_wgrad(const float* H, const float* G, f4 gW[4][4]) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    s4 bh[4], bl[4];
    for (int jb = 0; jb < 4; ++jb) {
        float bv[4];
        for (int e = 0; e < 4; ++e) bv[e] = G[(4 * g + e) * LDA + 16 * jb + r];
        split4(bv, bh[jb], bl[jb]);
    }
    for (int mb = 0; mb < 4; ++mb) {
        float av[4];
        for (int e = 0; e < 4; ++e) av[e] = H[(4 * g + e) * LDA + 16 * mb + r];
        s4 ah, al;
        split4(av, ah, al);
        for (int jb = 0; jb < 4; ++jb) gW[mb][jb] = mfma3(ah, al, bh[jb], bl[jb], gW[mb][jb]);
    }
}

// store a D-layout 16 x 64 result (+ bias) to LDS as pre-activation Z and activation H
__device__ __forceinline__ void store_act(const f4 acc[4], const float* __restrict__ bias, float* Z, float* Hh) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    for (int jb = 0; jb < 4; ++jb)
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * g + e, j = 16 * jb + r;
            const float z = acc[jb][e] + bias[j];
            Z[i * LDA + j] = z;
            Hh[i * LDA + j] = silu(z);
        }
}

// forward pass of one tile whose normalised inputs are in T.x; leaves Z1..3, H1..3, y in LDS
__device__ void forward_tile(const float* __restrict__ P, const WeightsLDS& W, TileLDS& T) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    // layer 1 (FP32): lane owns (i = 4g + e, j = 16 jb + r)
    for (int jb = 0; jb < 4; ++jb) {
        const int j = 16 * jb + r;
        const float w0 = P[OFF_W1 + j], w1 = P[OFF_W1 + NH + j], w2 = P[OFF_W1 + 2 * NH + j], b = P[OFF_B1 + j];
        for (int e = 0; e < 4; ++e) {
            const int i = 4 * g + e;
            const float z = b + T.x[i][0] * w0 + T.x[i][1] * w1 + T.x[i][2] * w2;
            T.z1[i * LDA + j] = z;
            T.h1[i * LDA + j] = silu(z);
        }
    }
    __syncthreads();
    f4 acc[4];
    gemm_fwd(T.h1, W.w2h, W.w2l, acc);
    store_act(acc, P + OFF_B2, T.z2, T.h2);
    __syncthreads();
    gemm_fwd(T.h2, W.w3h, W.w3l, acc);
    store_act(acc, P + OFF_B3, T.z3, T.h3);
    __syncthreads();
    // layer 4 (FP32): lanes 0..31 own (i = lane / 2, o = lane % 2)
    if (lane < 2 * TM) {
        const int i = lane >> 1, o = lane & 1;
        float y = P[OFF_B4 + o];
        for (int k = 0; k < NH; ++k) y += T.h3[i * LDA + k] * P[OFF_W4 + k * NOUT + o];
        T.y[i][o] = y;
    }
    __syncthreads();
}

__device__ __forceinline__ uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t i) {   // (0, 1)
    return ((splitmix(seed ^ splitmix(i)) >> 11) + 0.5) * 0x1p-53;
}

}  // namespace

// Exact coefficients in the surrogate coordinates (emission.cuh:159-170 normalised; FP64, bh_math
// elementary functions, the tracer's K2 table): (ln f_j, ln f_a) at ln X = lx, ln Theta = lth,
// ln sin theta_B = ls.
__device__ void exact_lnf_s(const float* __restrict__ k2, int k2_n, float k2_l0, float k2_dl, double lx, double lth,
                            double ls, double* lfj, double* lfa) {
    const double X = bhm::exp(lx), x3 = bhm::exp(lx * (1.0 / 3.0));
    const double brk = sqrt(X) + 1.887748625363387 * sqrt(x3);
    const double lk2 = (double)bhd::table_lerp(k2, k2_n, k2_l0, k2_dl, (float)lth);
    const double fj = LN_FJ_CONST + 2.0 * lth + ls - lk2 + 2.0 * bhm::log(brk) - x3;
    const double u0ln10 = lx + LN_2_9 + 2.0 * lth + ls;
    *lfj = fj;
    *lfa = fj - 0.6931471805599453 - lth - 2.0 * u0ln10;
}

// The same at u0 = log10(nu/nu_c), Theta, theta_B (sin theta_B >= 1e-3, as in the tracer).
__device__ void exact_lnf(const float* __restrict__ k2, int k2_n, float k2_l0, float k2_dl, double u0, double th,
                          double tb, double* lfj, double* lfa) {
    double s, c;
    bhm::sincos(tb, &s, &c);
    const double lth = bhm::log(th), ls = bhm::log(fmax(s, 1e-3));
    exact_lnf_s(k2, k2_n, k2_l0, k2_dl, u0 * LN10 - LN_2_9 - 2.0 * lth - ls, lth, ls, lfj, lfa);
}

namespace {

// training set: normalised inputs, raw (clamped) targets and loss weights
__global__ __launch_bounds__(256) void k_mlp_sample(uint64_t seed, int n, const float* __restrict__ k2, int k2_n,
                                                   float k2_l0, float k2_dl, Norm N, float hot_weight,
                                                   float* __restrict__ xs, float* __restrict__ ts,
                                                   float* __restrict__ ws) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double u0 = U0_LO + (U0_HI - U0_LO) * uniform01(seed, 3ull * i);
    const double lth = LTH_LO + (LTH_HI - LTH_LO) * uniform01(seed, 3ull * i + 1);
    const double tb = TB_LO + (TB_HI - TB_LO) * uniform01(seed, 3ull * i + 2);
    double sb, cb;
    bhm::sincos(tb, &sb, &cb);
    const double ls = bhm::log(fmax(sb, 1e-3));
    const double lx = u0 * LN10 - LN_2_9 - 2.0 * lth - ls;
    double fj, fa;
    exact_lnf_s(k2, k2_n, k2_l0, k2_dl, lx, lth, ls, &fj, &fa);
    const double sc[3] = {fmin(lx, LX_HI), lth, ls};
    for (int k = 0; k < 3; ++k) xs[3 * (size_t)i + k] = (float)((sc[k] - N.x_c[k]) / N.x_h[k]);
    ts[2 * (size_t)i] = (float)fmax(fj, FLOOR_J);
    ts[2 * (size_t)i + 1] = (float)fmax(fa, FLOOR_A);
    ws[i] = 1.f + hot_weight / (1.f + expf(-(float)(fj + 12.0)));
}

// Gradient workgroups: GRAD_WAVES waves share one copy of the split weights in LDS, and each wave
// takes GRAD_TILES 16-sample tiles with its own activation buffers (36 KB of weights + 35 KB per
// wave). The default batch of 8192 samples is then 256 workgroups, one per CU with two busy SIMDs.
// Measured per training step with one wave per workgroup: 4 tiles 104 us (half the CUs idle),
// 2 tiles 83 us, 1 tile 92 us (twice the weight loads and partial gradients to reduce).
constexpr int GRAD_WAVES = 2, GRAD_TILES = 1;
static_assert(GRAD_WAVES == 2, "k_mlp_grad combines the partial gradients of exactly two waves");

/* One training step's gradient over `count` samples starting at `first` (indices mod n): each
   workgroup (one wave) takes `tiles` consecutive 16-sample tiles and writes its partial gradient
   (8706 floats) and partial loss. Loss = sum_i w_i |y_i - t_i|^2 / count on normalised outputs. */
// The partial gradient held by one lane, visited in the parameter layout: f(index, value)
template <typename F>
__device__ __forceinline__ void for_partial(int lane, const float gW1[3], float gb1, float gb2, float gb3,
                                            const float gW4[2], float gb4, const f4 gW2[4][4], const f4 gW3[4][4],
                                            F f) {
    const int r = lane & 15, g = lane >> 4;
    for (int k = 0; k < 3; ++k) f(OFF_W1 + k * NH + lane, gW1[k]);
    f(OFF_B1 + lane, gb1);
    f(OFF_B2 + lane, gb2);
    f(OFF_B3 + lane, gb3);
    f(OFF_W4 + lane * NOUT, gW4[0]);
    f(OFF_W4 + lane * NOUT + 1, gW4[1]);
    if (lane < 2) f(OFF_B4 + lane, gb4);
    for (int mb = 0; mb < 4; ++mb)
        for (int jb = 0; jb < 4; ++jb)
            for (int e = 0; e < 4; ++e) {
                const int m = 16 * mb + 4 * g + e, j = 16 * jb + r;
                f(OFF_W2 + m * NH + j, gW2[mb][jb][e]);
                f(OFF_W3 + m * NH + j, gW3[mb][jb][e]);
            }
}

__global__ __launch_bounds__(64 * GRAD_WAVES) void k_mlp_grad(const float* __restrict__ P, Norm N, const float* __restrict__ xs,
                                                 const float* __restrict__ ts, const float* __restrict__ ws, int n,
                                                 int first, int count, int tiles, float* __restrict__ partial,
                                                 float* __restrict__ loss_partial) {
    __shared__ WeightsLDS W;
    __shared__ TileLDS TT[GRAD_WAVES];
    __shared__ float wave_loss[GRAD_WAVES];
    const int wave = threadIdx.x >> 6;
    TileLDS& T = TT[wave];
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    load_weights(P, W);
    __syncthreads();
    f4 gW2[4][4], gW3[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) gW2[a][b] = gW3[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    float gW1[3] = {0.f, 0.f, 0.f}, gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gW4[2] = {0.f, 0.f}, gb4 = 0.f, loss = 0.f;
    const float inv_count = 1.f / (float)count;
    for (int tl = 0; tl < tiles; ++tl) {
        const int base = ((blockIdx.x * GRAD_WAVES + wave) * tiles + tl) * TM;   // sample offset in the batch
        if (lane < TM) {
            const int s = lane;
            const bool in = base + s < count;
            const size_t idx = (size_t)((first + base + s) % n);
            for (int k = 0; k < 3; ++k) T.x[s][k] = in ? xs[3 * idx + k] : 0.f;
            T.t[s][0] = in ? (ts[2 * idx] - N.y_mu[0]) / N.y_sd[0] : 0.f;
            T.t[s][1] = in ? (ts[2 * idx + 1] - N.y_mu[1]) / N.y_sd[1] : 0.f;
            T.w[s] = in ? ws[idx] : 0.f;
        }
        __syncthreads();
        forward_tile(P, W, T);
        // loss gradient dY (normalised outputs), lanes 0..31
        if (lane < 2 * TM) {
            const int i = lane >> 1, o = lane & 1;
            const float d = T.y[i][o] - T.t[i][o];
            loss += T.w[i] * d * d * inv_count;
            T.y[i][o] = 2.f * T.w[i] * d * inv_count;   // reuse y as dY
        }
        __syncthreads();
        // layer 4: lane k owns W4[k][0..1]; dZ3[i][k] = (sum_o dY[i][o] W4[k][o]) silu'(Z3[i][k])
        {
            const int k = lane;
            const float w40 = P[OFF_W4 + k * NOUT], w41 = P[OFF_W4 + k * NOUT + 1];
            for (int i = 0; i < TM; ++i) {
                const float dy0 = T.y[i][0], dy1 = T.y[i][1];
                gW4[0] += T.h3[i * LDA + k] * dy0;
                gW4[1] += T.h3[i * LDA + k] * dy1;
                T.g0[i * LDA + k] = (dy0 * w40 + dy1 * w41) * dsilu(T.z3[i * LDA + k]);
            }
            if (lane < 2)
                for (int i = 0; i < TM; ++i) gb4 += T.y[i][lane];
        }
        __syncthreads();
        // layer 3
        gemm_wgrad(T.h2, T.g0, gW3);
        for (int i = 0; i < TM; ++i) gb3 += T.g0[i * LDA + lane];
        f4 acc[4];
        gemm_bwd(T.g0, W.w3h, W.w3l, acc);
        for (int kb = 0; kb < 4; ++kb)
            for (int e = 0; e < 4; ++e) {
                const int i = 4 * g + e, k = 16 * kb + r;
                T.g1[i * LDA + k] = acc[kb][e] * dsilu(T.z2[i * LDA + k]);
            }
        __syncthreads();
        // layer 2
        gemm_wgrad(T.h1, T.g1, gW2);
        for (int i = 0; i < TM; ++i) gb2 += T.g1[i * LDA + lane];
        gemm_bwd(T.g1, W.w2h, W.w2l, acc);
        __syncthreads();   // every lane has read g0's predecessor data; g0 is reused below
        for (int kb = 0; kb < 4; ++kb)
            for (int e = 0; e < 4; ++e) {
                const int i = 4 * g + e, k = 16 * kb + r;
                T.g0[i * LDA + k] = acc[kb][e] * dsilu(T.z1[i * LDA + k]);
            }
        __syncthreads();
        // layer 1: lane j owns W1[0..2][j], b1[j]
        for (int i = 0; i < TM; ++i) {
            const float d = T.g0[i * LDA + lane];
            gW1[0] += T.x[i][0] * d;
            gW1[1] += T.x[i][1] * d;
            gW1[2] += T.x[i][2] * d;
            gb1 += d;
        }
        __syncthreads();
    }
    // the workgroup's partial gradient: wave 1 leaves its sums in its own (no longer used) tile
    // buffer, wave 0 adds them to its own and writes the result, always in that order
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    if (lane == 0) wave_loss[wave] = loss;
    __syncthreads();
    float* red = reinterpret_cast<float*>(&TT[1]);
    static_assert(sizeof(TileLDS) >= sizeof(float) * NPARAM, "a tile buffer holds a partial gradient");
    if (wave == 1) for_partial(lane, gW1, gb1, gb2, gb3, gW4, gb4, gW2, gW3, [&](int i, float v) { red[i] = v; });
    __syncthreads();
    if (wave == 0) {
        float* out = partial + (size_t)blockIdx.x * NPARAM;
        for_partial(lane, gW1, gb1, gb2, gb3, gW4, gb4, gW2, gW3, [&](int i, float v) { out[i] = v + red[i]; });
        if (lane == 0) loss_partial[blockIdx.x] = wave_loss[0] + wave_loss[1];
    }
}

// grad[p] = sum over workgroups of partial[w][p]; loss likewise. A workgroup takes 64 parameters; its
// eight waves each sum an eighth of the workgroups' partials (w = q, q + 8, ...) in eight running
// sums (eight independent loads in flight), and the eighths are added through LDS. Every sum is
// taken in a fixed order, so the result does not depend on timing.
// With P set, the same lane then takes the Adam step of its parameter (bias-corrected moments).
struct AdamStep {
    float* P;   // nullptr: gradient only
    float* m;
    float* v;
    float lr, b1, b2, eps, bc1, bc2;
};

constexpr int RED_PARAMS = 64, RED_SPLIT = 8, RED_BLOCK = RED_PARAMS * RED_SPLIT;
__global__ __launch_bounds__(RED_BLOCK) void k_mlp_reduce(const float* __restrict__ partial, int nw,
                                                         const float* __restrict__ loss_partial,
                                                         float* __restrict__ grad, float* __restrict__ loss,
                                                         AdamStep A) {
    __shared__ float part_sum[RED_SPLIT][RED_PARAMS];
    static_assert(RED_SPLIT == 8, "the final sum below adds eight parts");
    const int lp = threadIdx.x % RED_PARAMS, qt = threadIdx.x / RED_PARAMS;
    const int p = blockIdx.x * RED_PARAMS + lp;
    float total = 0.f;
    if (p < NPARAM) {
        constexpr size_t STEP = (size_t)RED_SPLIT * NPARAM;   // between this group's rows
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f, s6 = 0.f, s7 = 0.f, sr = 0.f;
        const float* q = partial + (size_t)qt * NPARAM + p;
        int w = qt;
        for (; w + 7 * RED_SPLIT < nw; w += 8 * RED_SPLIT, q += 8 * STEP) {
            s0 += q[0];
            s1 += q[STEP];
            s2 += q[2 * STEP];
            s3 += q[3 * STEP];
            s4 += q[4 * STEP];
            s5 += q[5 * STEP];
            s6 += q[6 * STEP];
            s7 += q[7 * STEP];
        }
        for (; w < nw; w += RED_SPLIT, q += STEP) sr += q[0];
        total = (((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7))) + sr;
    }
    part_sum[qt][lp] = total;
    __syncthreads();
    if (qt == 0 && p < NPARAM) {
        const float(&q)[RED_SPLIT][RED_PARAMS] = part_sum;
        const float gr = ((q[0][lp] + q[1][lp]) + (q[2][lp] + q[3][lp])) + ((q[4][lp] + q[5][lp]) + (q[6][lp] + q[7][lp]));
        grad[p] = gr;
        if (A.P) {
            const float mm = A.b1 * A.m[p] + (1.f - A.b1) * gr, vv = A.b2 * A.v[p] + (1.f - A.b2) * gr * gr;
            A.m[p] = mm;
            A.v[p] = vv;
            A.P[p] -= A.lr * (mm / A.bc1) / (sqrtf(vv / A.bc2) + A.eps);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 64) {   // the loss: one wave, strided sums and a fixed tree
        float sl = 0.f;
        for (int v = threadIdx.x; v < nw; v += 64) sl += loss_partial[v];
        for (int off = 32; off > 0; off >>= 1) sl += __shfl_down(sl, off, 64);
        if (threadIdx.x == 0) *loss = sl;
    }
}


/* Inference: n samples of raw inputs (u0, Theta, theta_B) or, with lattice != 0, the nodes of the
   (n0 x n1 x n2) lattice over (u0, log10 Theta, theta_B in [lo2, hi2]); writes (ln f_j, ln f_a). */
struct Lattice {
    int n0, n1, n2;
    float lo0, lo1, lo2, d0, d1, d2;
};
__global__ __launch_bounds__(64) void k_mlp_infer(const float* __restrict__ P, Norm N, const float* __restrict__ in,
                                                  int n, int lattice, Lattice L, float2* __restrict__ out) {
    __shared__ WeightsLDS W;
    __shared__ TileLDS T;
    const int lane = threadIdx.x & 63;
    load_weights(P, W);
    __syncthreads();
    const int ntiles = (n + TM - 1) / TM;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        if (lane < TM) {
            const int s = tile * TM + lane;
            float sc[3] = {0.f, 0.f, 0.f};   // surrogate coordinates
            if (s < n) {
                if (lattice) {
                    const int i0 = s % L.n0, i1 = (s / L.n0) % L.n1, i2 = s / (L.n0 * L.n1);
                    sc[0] = L.lo0 + i0 * L.d0;
                    sc[1] = L.lo1 + i1 * L.d1;
                    sc[2] = L.lo2 + i2 * L.d2;
                } else {
                    const float u0 = in[3 * (size_t)s], lth = logf(in[3 * (size_t)s + 1]);
                    const float ls = logf(fmaxf(sinf(in[3 * (size_t)s + 2]), 1e-3f));
                    sc[0] = fminf(u0 * (float)LN10 - (float)LN_2_9 - 2.f * lth - ls, (float)LX_HI);
                    sc[1] = lth;
                    sc[2] = ls;
                }
            }
            for (int k = 0; k < 3; ++k) T.x[lane][k] = (sc[k] - N.x_c[k]) / N.x_h[k];
        }
        __syncthreads();
        forward_tile(P, W, T);
        if (lane < TM) {
            const int s = tile * TM + lane;
            if (s < n) out[s] = make_float2(T.y[lane][0] * N.y_sd[0] + N.y_mu[0], T.y[lane][1] * N.y_sd[1] + N.y_mu[1]);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_mlp_exact(const float* __restrict__ in, int n, const float* __restrict__ k2,
                                                   int k2_n, float k2_l0, float k2_dl, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double fj, fa;
    exact_lnf(k2, k2_n, k2_l0, k2_dl, in[3 * (size_t)i], in[3 * (size_t)i + 1], in[3 * (size_t)i + 2], &fj, &fa);
    out[2 * (size_t)i] = fj;
    out[2 * (size_t)i + 1] = fa;
}

/* Judge: for every lattice cell, the exact coefficients at 4 interior points against the trilinear
   interpolation of the tabulated network; a cell fails when |d ln f_j| > tol where ln f_j > floor
   (or the same for ln f_a). */
__global__ __launch_bounds__(256) void k_mlp_judge(const float2* __restrict__ tab, Lattice L, const float* __restrict__ k2,
                                                   int k2_n, float k2_l0, float k2_dl, float tol, float floor_j,
                                                   float floor_a, unsigned char* __restrict__ bad,
                                                   unsigned int* __restrict__ counts) {
    const int c0 = L.n0 - 1, c1 = L.n1 - 1, c2 = L.n2 - 1;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= c0 * c1 * c2) return;
    const int i0 = cell % c0, i1 = (cell / c0) % c1, i2 = cell / (c0 * c1);
    const float2* b = tab + ((size_t)i2 * L.n1 + i1) * L.n0 + i0;
    const size_t s1 = L.n0, s2 = (size_t)L.n0 * L.n1;
    const float2 v[8] = {b[0], b[1], b[s1], b[s1 + 1], b[s2], b[s2 + 1], b[s2 + s1], b[s2 + s1 + 1]};
    const float fr[4][3] = {{0.5f, 0.5f, 0.5f}, {0.25f, 0.25f, 0.75f}, {0.75f, 0.75f, 0.25f}, {0.25f, 0.75f, 0.5f}};
    bool fail = false, hot = false;
    for (int p = 0; p < 4; ++p) {
        const float f0 = fr[p][0], f1 = fr[p][1], f2 = fr[p][2];
        float iv[2];
        for (int o = 0; o < 2; ++o) {
            auto at = [&](int q) { return o == 0 ? v[q].x : v[q].y; };
            float c00 = at(0) + (at(1) - at(0)) * f0, c10 = at(2) + (at(3) - at(2)) * f0;
            float c01 = at(4) + (at(5) - at(4)) * f0, c11 = at(6) + (at(7) - at(6)) * f0;
            float c0v = c00 + (c10 - c00) * f1, c1v = c01 + (c11 - c01) * f1;
            iv[o] = c0v + (c1v - c0v) * f2;
        }
        const double lx = L.lo0 + (i0 + f0) * L.d0, lth = L.lo1 + (i1 + f1) * L.d1, ls = L.lo2 + (i2 + f2) * L.d2;
        double fj, fa;
        exact_lnf_s(k2, k2_n, k2_l0, k2_dl, lx, lth, ls, &fj, &fa);
        if (fj > floor_j) {
            hot = true;
            if (fabs(iv[0] - fj) > tol) fail = true;
        }
        if (fa > floor_a && fabs(iv[1] - fa) > tol) fail = true;
    }
    bad[cell] = fail ? 1 : 0;
    if (hot) atomicAdd(&counts[0], 1u);
    if (hot && fail) atomicAdd(&counts[1], 1u);
}

}  // namespace

#define MCK(ctx, x)                                                                                        \
    do {                                                                                                   \
        hipError_t e_ = (x);                                                                               \
        if (e_ != hipSuccess) {                                                                            \
            (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);                                   \
            return BH_ERR_HIP;                                                                             \
        }                                                                                                  \
    } while (0)

namespace {

Norm default_norm() {
    Norm N;
    const double lo[3] = {LX_LO, LTH_LO, LS_LO}, hi[3] = {LX_HI, LTH_HI, LS_HI};
    for (int k = 0; k < 3; ++k) {
        N.x_c[k] = (float)(0.5 * (lo[k] + hi[k]));
        N.x_h[k] = (float)(0.5 * (hi[k] - lo[k]));
    }
    N.y_mu[0] = N.y_mu[1] = 0.f;
    N.y_sd[0] = N.y_sd[1] = 1.f;
    return N;
}

void norm_to_floats(const Norm& N, float* v) {
    for (int k = 0; k < 3; ++k) {
        v[k] = N.x_c[k];
        v[3 + k] = N.x_h[k];
    }
    v[6] = N.y_mu[0];
    v[7] = N.y_mu[1];
    v[8] = N.y_sd[0];
    v[9] = N.y_sd[1];
}

Norm norm_from_floats(const float* v) {
    Norm N;
    for (int k = 0; k < 3; ++k) {
        N.x_c[k] = v[k];
        N.x_h[k] = v[3 + k];
    }
    N.y_mu[0] = v[6];
    N.y_mu[1] = v[7];
    N.y_sd[0] = v[8];
    N.y_sd[1] = v[9];
    return N;
}

int ensure_params(bh_ctx* c) {
    if (c->d_mlp) return BH_OK;
    MCK(c, hipMalloc((void**)&c->d_mlp, sizeof(float) * NPARAM));
    MCK(c, hipMalloc((void**)&c->d_adam_m, sizeof(float) * NPARAM));
    MCK(c, hipMalloc((void**)&c->d_adam_v, sizeof(float) * NPARAM));
    return BH_OK;
}

// He-style uniform initialisation (host, seeded)
void init_params(uint64_t seed, std::vector<float>* p) {
    p->assign(NPARAM, 0.f);
    uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        return (double)(s >> 11) * 0x1p-53 * 2.0 - 1.0;
    };
    auto fill = [&](int off, int fan_in, int count) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) (*p)[off + i] = (float)(a * next());
    };
    fill(OFF_W1, NIN, NIN * NH);
    fill(OFF_W2, NH, NH * NH);
    fill(OFF_W3, NH, NH * NH);
    fill(OFF_W4, NH, NH * NOUT);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int bh_mlp_train(bh_ctx* c, const bh_mlp_train_cfg* cfg_in, bh_mlp_report* rep) {
    if (!c) return BH_ERR_ARG;
    if (!c->d_k2) {
        c->err = "the surrogate's exact routine needs the K2 table (radio scene)";
        return BH_ERR_ARG;
    }
    bh_mlp_train_cfg cfg{};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.n_samples <= 0) cfg.n_samples = 1 << 22;
    if (cfg.batch <= 0) cfg.batch = 8192;
    if (cfg.epochs <= 0) cfg.epochs = 200;
    if (!(cfg.lr > 0)) cfg.lr = 2e-3f;
    if (!(cfg.lr_min > 0)) cfg.lr_min = 2e-5f;
    if (cfg.seed == 0) cfg.seed = 1;
    if (!(cfg.hot_weight >= 0)) cfg.hot_weight = 20.f;
    if (cfg.batch % TM != 0 || cfg.batch > cfg.n_samples) {
        c->err = "batch must be a multiple of 16 and at most n_samples";
        return BH_ERR_ARG;
    }
    MCK(c, hipSetDevice(c->device));
    int rc = ensure_params(c);
    if (rc) return rc;
    const int n = cfg.n_samples;
    float *xs = nullptr, *ts = nullptr, *ws = nullptr, *partial = nullptr, *lpart = nullptr, *grad = nullptr,
          *loss = nullptr;
    const int tiles = GRAD_TILES;
    const int nwg = (cfg.batch / TM + tiles * GRAD_WAVES - 1) / (tiles * GRAD_WAVES);
    auto cleanup = [&]() {
        for (float* p : {xs, ts, ws, partial, lpart, grad, loss})
            if (p) (void)hipFree(p);
    };
    double t0 = now_ms();
    if (hipMalloc((void**)&xs, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&ts, sizeof(float) * 2 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&ws, sizeof(float) * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&partial, sizeof(float) * NPARAM * (size_t)nwg) != hipSuccess ||
        hipMalloc((void**)&lpart, sizeof(float) * nwg) != hipSuccess ||
        hipMalloc((void**)&grad, sizeof(float) * NPARAM) != hipSuccess ||
        hipMalloc((void**)&loss, sizeof(float)) != hipSuccess) {
        cleanup();
        c->err = "surrogate training buffers: hipMalloc failed";
        return BH_ERR_HIP;
    }
    Norm N = default_norm();
    hipLaunchKernelGGL(k_mlp_sample, dim3((n + 255) / 256), dim3(256), 0, c->st, cfg.seed, n, c->d_k2, c->k2_n,
                       c->k2_l0, c->k2_dl, N, cfg.hot_weight, xs, ts, ws);
    // output normalisation from the samples
    std::vector<float> th((size_t)2 * n);
    if (hipMemcpyAsync(th.data(), ts, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess) {
        cleanup();
        c->err = "surrogate sampler failed";
        return BH_ERR_HIP;
    }
    for (int o = 0; o < 2; ++o) {
        double s = 0, s2 = 0;
        for (int i = 0; i < n; ++i) {
            s += th[2 * (size_t)i + o];
            s2 += (double)th[2 * (size_t)i + o] * th[2 * (size_t)i + o];
        }
        const double mu = s / n;
        N.y_mu[o] = (float)mu;
        N.y_sd[o] = (float)std::sqrt(std::max(s2 / n - mu * mu, 1e-12));
    }
    double t1 = now_ms();
    std::vector<float> p0;
    init_params(cfg.seed, &p0);
    std::vector<float> zeros(NPARAM, 0.f);
    if (hipMemcpyAsync(c->d_mlp, p0.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemcpyAsync(c->d_adam_m, zeros.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemcpyAsync(c->d_adam_v, zeros.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->st) != hipSuccess) {
        cleanup();
        c->err = "surrogate parameter upload failed";
        return BH_ERR_HIP;
    }
    const int steps_per_epoch = n / cfg.batch;
    const int total = steps_per_epoch * cfg.epochs;
    float first_loss = -1.f, last_loss = -1.f;
    for (int step = 0; step < total; ++step) {
        const int first = (int)(((long long)step * cfg.batch) % n);
        hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64 * GRAD_WAVES), 0, c->st, c->d_mlp, N, xs, ts, ws, n, first,
                           cfg.batch,
                           tiles, partial, lpart);
        const double prog = (double)step / std::max(total - 1, 1);
        const float lr = (float)(cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1.0 + std::cos(3.141592653589793 * prog)));
        const float bc1 = (float)(1.0 - std::pow(0.9, step + 1)), bc2 = (float)(1.0 - std::pow(0.999, step + 1));
        const AdamStep A{c->d_mlp, c->d_adam_m, c->d_adam_v, lr, 0.9f, 0.999f, 1e-8f, bc1, bc2};
        hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + RED_PARAMS - 1) / RED_PARAMS), dim3(RED_BLOCK), 0, c->st,
                           partial, nwg, lpart, grad, loss, A);
        if (step == 0 || step == total - 1) {
            float l = 0;
            if (hipMemcpyAsync(&l, loss, sizeof(float), hipMemcpyDeviceToHost, c->st) != hipSuccess ||
                hipStreamSynchronize(c->st) != hipSuccess) {
                cleanup();
                c->err = "surrogate training step failed";
                return BH_ERR_HIP;
            }
            if (step == 0) first_loss = l;
            else last_loss = l;
        }
    }
    if (hipStreamSynchronize(c->st) != hipSuccess || hipGetLastError() != hipSuccess) {
        cleanup();
        c->err = "surrogate training failed";
        return BH_ERR_HIP;
    }
    double t2 = now_ms();
    cleanup();
    c->mlp_norm = N;
    c->mlp_trained = true;
    c->stab_ready = false;
    if (rep) {
        rep->ms_sample = t1 - t0;
        rep->ms_train = t2 - t1;
        rep->steps = total;
        rep->loss_first = first_loss;
        rep->loss_last = last_loss;
        // 3 MFMA products per hidden GEMM, forward + 2 backward GEMMs per hidden layer
        rep->train_flop = (double)total * cfg.batch * 2.0 * (2.0 * NH * NH) * 3.0 * 3.0;
    }
    return BH_OK;
}

int bh_mlp_get(bh_ctx* c, float* params, float* norm) {
    if (!c || !params) return BH_ERR_ARG;
    if (!c->mlp_trained) {
        c->err = "no surrogate parameters";
        return BH_ERR_ARG;
    }
    MCK(c, hipSetDevice(c->device));
    MCK(c, hipMemcpy(params, c->d_mlp, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    if (norm) norm_to_floats(c->mlp_norm, norm);
    return BH_OK;
}

int bh_mlp_set(bh_ctx* c, const float* params, const float* norm) {
    if (!c || !params || !norm) return BH_ERR_ARG;
    MCK(c, hipSetDevice(c->device));
    int rc = ensure_params(c);
    if (rc) return rc;
    MCK(c, hipMemcpy(c->d_mlp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    c->mlp_norm = norm_from_floats(norm);
    c->mlp_trained = true;
    c->stab_ready = false;
    return BH_OK;
}

int bh_mlp_infer(bh_ctx* c, const float* in, int n, float* out) {
    if (!c || !in || !out || n < 0) return BH_ERR_ARG;
    if (!c->mlp_trained) {
        c->err = "no surrogate parameters";
        return BH_ERR_ARG;
    }
    if (n == 0) return BH_OK;
    MCK(c, hipSetDevice(c->device));
    float* din = nullptr;
    float2* dout = nullptr;
    int rc = BH_OK;
    if (hipMalloc((void**)&din, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&dout, sizeof(float2) * (size_t)n) != hipSuccess ||
        hipMemcpy(din, in, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = BH_ERR_HIP;
    } else {
        const int ntiles = (n + TM - 1) / TM;
        Lattice L{};
        hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(ntiles, 4096)), dim3(64), 0, c->st, c->d_mlp, c->mlp_norm, din,
                           n, 0, L, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, sizeof(float2) * (size_t)n, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess)
            rc = BH_ERR_HIP;
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (rc) c->err = "surrogate inference failed";
    return rc;
}

int bh_mlp_exact(bh_ctx* c, const float* in, int n, double* out) {
    if (!c || !in || !out || n < 0) return BH_ERR_ARG;
    if (!c->d_k2) {
        c->err = "the exact routine needs the K2 table";
        return BH_ERR_ARG;
    }
    if (n == 0) return BH_OK;
    MCK(c, hipSetDevice(c->device));
    float* din = nullptr;
    double* dout = nullptr;
    int rc = BH_OK;
    if (hipMalloc((void**)&din, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&dout, sizeof(double) * 2 * (size_t)n) != hipSuccess ||
        hipMemcpy(din, in, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = BH_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_mlp_exact, dim3((n + 255) / 256), dim3(256), 0, c->st, din, n, c->d_k2, c->k2_n, c->k2_l0,
                           c->k2_dl, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
            hipStreamSynchronize(c->st) != hipSuccess)
            rc = BH_ERR_HIP;
    }
    if (din) (void)hipFree(din);
    if (dout) (void)hipFree(dout);
    if (rc) c->err = "exact routine failed";
    return rc;
}

int bh_mlp_grad(bh_ctx* c, const float* params, const float* norm, const float* xn, const float* tn, const float* w,
                int n, float* loss_out, float* grad_out) {
    if (!c || !params || !norm || !xn || !tn || !w || !grad_out || n <= 0 || n % TM != 0) return BH_ERR_ARG;
    MCK(c, hipSetDevice(c->device));
    const Norm N = norm_from_floats(norm);
    const int tiles = GRAD_TILES, nwg = (n / TM + tiles * GRAD_WAVES - 1) / (tiles * GRAD_WAVES);
    float *dp = nullptr, *dx = nullptr, *dt = nullptr, *dw = nullptr, *part = nullptr, *lp = nullptr, *g = nullptr,
          *l = nullptr;
    int rc = BH_OK;
    if (hipMalloc((void**)&dp, sizeof(float) * NPARAM) != hipSuccess ||
        hipMalloc((void**)&dx, sizeof(float) * 3 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&dt, sizeof(float) * 2 * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&dw, sizeof(float) * (size_t)n) != hipSuccess ||
        hipMalloc((void**)&part, sizeof(float) * NPARAM * (size_t)nwg) != hipSuccess ||
        hipMalloc((void**)&lp, sizeof(float) * nwg) != hipSuccess ||
        hipMalloc((void**)&g, sizeof(float) * NPARAM) != hipSuccess || hipMalloc((void**)&l, sizeof(float)) != hipSuccess ||
        hipMemcpy(dp, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, xn, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, tn, sizeof(float) * 2 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dw, w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        rc = BH_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_mlp_grad, dim3(nwg), dim3(64 * GRAD_WAVES), 0, c->st, dp, N, dx, dt, dw, n, 0, n, tiles,
                           part, lp);
        hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + RED_PARAMS - 1) / RED_PARAMS), dim3(RED_BLOCK), 0, c->st, part,
                           nwg, lp, g, l, AdamStep{});
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(grad_out, g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
            (loss_out && hipMemcpyAsync(loss_out, l, sizeof(float), hipMemcpyDeviceToHost, c->st) != hipSuccess) ||
            hipStreamSynchronize(c->st) != hipSuccess)
            rc = BH_ERR_HIP;
    }
    for (float* p : {dp, dx, dt, dw, part, lp, g, l})
        if (p) (void)hipFree(p);
    if (rc) c->err = "surrogate gradient failed";
    return rc;
}

int bh_mlp_bake(bh_ctx* c, int n0, int n1, int n2, float judge_tol, bh_mlp_bake_report* rep) {
    if (!c) return BH_ERR_ARG;
    if (!c->mlp_trained || !c->d_k2) {
        c->err = "baking needs trained surrogate parameters and the K2 table";
        return BH_ERR_ARG;
    }
    if (n0 <= 0) n0 = 1024;
    if (n1 <= 0) n1 = 128;
    if (n2 <= 0) n2 = 8;
    if (!(judge_tol > 0)) judge_tol = 2e-3f;
    if (n0 < 2 || n1 < 2 || n2 < 2 || (size_t)n0 * n1 * n2 > (size_t)1 << 24) {
        c->err = "bad surrogate lattice size";
        return BH_ERR_ARG;
    }
    MCK(c, hipSetDevice(c->device));
    if (c->d_stab) (void)hipFree(c->d_stab);
    if (c->d_sbad) (void)hipFree(c->d_sbad);
    c->d_stab = nullptr;
    c->d_sbad = nullptr;
    c->stab_ready = false;
    const size_t nodes = (size_t)n0 * n1 * n2, cells = (size_t)(n0 - 1) * (n1 - 1) * (n2 - 1);
    MCK(c, hipMalloc((void**)&c->d_stab, sizeof(float2) * nodes));
    MCK(c, hipMalloc((void**)&c->d_sbad, cells));
    unsigned int* counts = nullptr;
    MCK(c, hipMalloc((void**)&counts, 2 * sizeof(unsigned int)));
    // lattice over the surrogate coordinates (ln X, ln Theta, ln sin theta_B)
    Lattice L;
    L.n0 = n0;
    L.n1 = n1;
    L.n2 = n2;
    L.lo0 = (float)LX_LO;
    L.d0 = (float)((LX_HI - LX_LO) / (n0 - 1));
    L.lo1 = (float)LTH_LO;
    L.d1 = (float)((LTH_HI - LTH_LO) / (n1 - 1));
    L.lo2 = (float)LS_LO;
    L.d2 = (float)((LS_HI - LS_LO) / (n2 - 1));
    double t0 = now_ms();
    const int ntiles = (int)((nodes + TM - 1) / TM);
    hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(ntiles, 8192)), dim3(64), 0, c->st, c->d_mlp, c->mlp_norm, nullptr,
                       (int)nodes, 1, L, c->d_stab);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, c->st);
    (void)hipMemsetAsync(counts, 0, 2 * sizeof(unsigned int), c->st);
    hipLaunchKernelGGL(k_mlp_judge, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c->st, c->d_stab, L, c->d_k2,
                       c->k2_n, c->k2_l0, c->k2_dl, judge_tol, (float)(FLOOR_J + 10.0), (float)(FLOOR_A + 10.0),
                       c->d_sbad, counts);
    (void)hipEventRecord(e1, c->st);
    unsigned int hc[2] = {0, 0};
    int rc = BH_OK;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
        rc = BH_ERR_HIP;
    float ms_judge = 0.f;
    (void)hipEventElapsedTime(&ms_judge, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(counts);
    if (rc) {
        c->err = "surrogate bake failed";
        return rc;
    }
    SurrogateTable& T = c->stab;
    T.tab = (const float*)c->d_stab;
    T.bad = c->d_sbad;
    T.n0 = n0;
    T.n1 = n1;
    T.n2 = n2;
    T.lo0 = L.lo0;
    T.lo1 = L.lo1;
    T.lo2 = L.lo2;
    T.inv_d0 = 1.f / L.d0;
    T.inv_d1 = 1.f / L.d1;
    T.inv_d2 = 1.f / L.d2;
    T.hi0 = (float)LX_HI;
    c->stab_ready = true;
    if (rep) {
        rep->ms_bake = now_ms() - t0;
        rep->ms_judge = ms_judge;
        rep->cells = (long long)cells;
        rep->hot_cells = hc[0];
        rep->failed_hot_cells = hc[1];
    }
    return BH_OK;
}

}  // extern "C"
