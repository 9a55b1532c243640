/* bh_math.h - FP64 elementary functions that give bit-identical results on the x86-64 host
   (g++, SSE2) and on gfx950 (hipcc), provided both sides compile with -ffp-contract=off.

   Why: the adaptive integrator branches on values that pass through pow/exp/log/atan2 every step
   (step-size controller, swept angle, emission). Near the photon ring the geodesic flow is
   unstable (Lyapunov factor ~e^pi per half orbit), so a one-ulp difference between two libm
   implementations grows into different step counts, winding counts and even fates. Routing every
   per-step transcendental through this header makes the GPU tracer and the CPU oracle agree bit
   for bit; accuracy against glibc is <= 3 ulp (pinned by tests/test_bh_math.py over 10^6 inputs per
   function), and x^-0.2 is correctly rounded.

   Methods (textbook reductions; the polynomials are near-minimax fits of this project, least squares
   on Chebyshev nodes with Lawson reweighting in 120-bit arithmetic, coefficients rounded to double;
   the fit error is below 2^-58 relative in every case). Polynomials are evaluated in Horner form
   with explicit fused multiply-adds: an FMA is correctly rounded on both sides (v_fma_f64 on the
   device, the C library's fma on the host), so the results stay bit-identical while each Horner
   step costs one instruction instead of a separate multiply and add.
   - exp, expm1: Cody-Waite reduction x = k ln2 + r with a 32-bit ln2 head, |r| <= ln2/2;
     e^r - 1 = r + r^2 q(r) with q of degree 10; expm1 recombines 2^k expm1(r) + (2^k-1).
   - log: x = m 2^e, m in [sqrt(1/2), sqrt(2)], f = m-1, s = f/(2+f), log(1+f) = f - s (f - R) with
     R = 2 atanh(s)/s - 2 = z P(z), z = s^2, P of degree 6.
   - cbrt: m in [1,2) and the exponent split by 3; two Newton steps w <- w (4 - m w^3)/3 for
     m^{-1/3} from a cubic least-squares seed (5e-4), cbrt(m) = m w^2, then one Newton polish.
   - sincos: Cody-Waite reduction by pi/2 with a three-part constant (33+33+53 bits);
     sin r = r + r z S(z) (S of degree 5), cos r = 1 + z C(z) (C of degree 6) on |r| <= pi/4.
   - atan2/acos: reciprocal and pi/6 (Hart) reductions to |t| <= tan(pi/12),
     atan t = t + t z A(z) with A of degree 8; acos(x) = atan2(sqrt((1-x)(1+x)), x),
     asin(x) = atan2(x, sqrt((1-x)(1+x))).
   - x^-0.2 (the step controller): exp(-0.2 log x) refined by one Newton step on x y^5 = 1 with an
     FMA double-double residual; correctly rounded on [1e-300, 1e300].
   Subnormal results of exp are flushed to zero (deterministically on both sides). */
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define BHM_HD __host__ __device__ inline
#else
#define BHM_HD inline
#endif

namespace bhm {

BHM_HD uint64_t dbits(double x) { return __builtin_bit_cast(uint64_t, x); }
BHM_HD double dfrom(uint64_t b) { return __builtin_bit_cast(double, b); }
BHM_HD double pow2i(int k) { return dfrom((uint64_t)(k + 1023) << 52); }   // -1022 <= k <= 1023
BHM_HD double fabs_(double x) { return dfrom(dbits(x) & 0x7fffffffffffffffull); }
BHM_HD double floor_(double x) { return __builtin_floor(x); }
BHM_HD bool isnan_(double x) { return !(x == x); }
BHM_HD bool isinf_(double x) { return fabs_(x) == __builtin_inf(); }

constexpr double LN2_HI = 0x1.62e42fee00000p-1;
constexpr double LN2_LO = 0x1.a39ef35793c76p-33;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PI_ = 0x1.921fb54442d18p+1;
constexpr double PI_2 = 0x1.921fb54442d18p+0;
constexpr double PI_6 = 0x1.0c152382d7366p-1;
constexpr double SQRT3 = 0x1.bb67ae8584caap+0;
constexpr double TAN_PI_12 = 0x1.126145e9ecd56p-2;
constexpr double SQRT2 = 0x1.6a09e667f3bcdp+0;
constexpr double CBRT2 = 0x1.428a2f98d728bp+0;
constexpr double CBRT4 = 0x1.965fea53d6e3dp+0;

BHM_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// One Horner step p x + c with c a constant coefficient. gfx950 has no 64-bit literal operand, and
// the compiler turns fma(p, x, literal) into v_fmac_f64 with the coefficient as the tied addend,
// copied into a VGPR pair first (two v_mov_b32 per step). Here the coefficient sits in a scalar
// register pair (one VOP3 operand may be scalar) and the step is one v_fma_f64: one vector
// instruction instead of three. Same correctly rounded FMA as the host's.
BHM_HD double horner_(double p, double x, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    double d;
    __asm__("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(p), "v"(x), "s"(c));
    return d;
#else
    return __builtin_fma(p, x, c);
#endif
}

// (e^r - 1 - r) / r^2 for |r| <= ln2/2, degree 10 (fit error 2.6e-18)
BHM_HD double expm1_core_q(double r) {
    double p = 0x1.1f8b3e103c1fep-29;
    p = horner_(p, r, 0x1.af4dde3ddccc9p-26);
    p = horner_(p, r, 0x1.27e4ccd670e98p-22);
    p = horner_(p, r, 0x1.71de0236f22adp-19);
    p = horner_(p, r, 0x1.a01a01acf7424p-16);
    p = horner_(p, r, 0x1.a01a01abe651ep-13);
    p = horner_(p, r, 0x1.6c16c16c15206p-10);
    p = horner_(p, r, 0x1.11111111100dfp-7);
    p = horner_(p, r, 0x1.5555555555558p-5);
    p = horner_(p, r, 0x1.5555555555557p-3);
    p = fma_(p, r, 0.5);
    return p;
}

BHM_HD double exp(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -708.3964185322641) return 0.0;   // result would be subnormal: flushed
    double kd = floor_(fma_(x, INV_LN2, 0.5));
    int k = (int)kd;
    double r = fma_(-kd, LN2_LO, fma_(-kd, LN2_HI, x));
    double e = 1.0 + fma_(r * r, expm1_core_q(r), r);
    if (k > 1023) return e * 2.0 * pow2i(k - 1);
    if (k < -1021) return (e * pow2i(k + 54)) * 0x1p-54;
    return e * pow2i(k);
}

BHM_HD double expm1(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -38.0) return -1.0;   // e^x < 2^-54
    if (fabs_(x) <= 0.34657359027997264) return fma_(x * x, expm1_core_q(x), x);
    double kd = floor_(fma_(x, INV_LN2, 0.5));
    int k = (int)kd;
    double r = fma_(-kd, LN2_LO, fma_(-kd, LN2_HI, x));
    double em = fma_(r * r, expm1_core_q(r), r);
    if (k > 56) {
        double e = em + 1.0;
        if (k > 1023) return e * 2.0 * pow2i(k - 1);
        return e * pow2i(k);
    }
    double t = pow2i(k);
    return fma_(t, em, t - 1.0);
}

BHM_HD double log(double x) {
    if (isnan_(x)) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (isinf_(x)) return x;
    uint64_t b = dbits(x);
    int e = (int)(b >> 52) - 1023;
    if (e == -1023) {   // subnormal input
        b = dbits(x * 0x1p54);
        e = (int)(b >> 52) - 1023 - 54;
    }
    double m = dfrom((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    // R = 2 atanh(s) / s - 2 = z P(z), P of degree 6 (fit error of R 3.6e-18)
    double P = 0x1.2f056aef6994ap-3;
    P = horner_(P, z, 0x1.39a1b9ea2f5d1p-3);
    P = horner_(P, z, 0x1.74663ef13705bp-3);
    P = horner_(P, z, 0x1.c71c52092cd21p-3);
    P = horner_(P, z, 0x1.24924941f169ep-2);
    P = horner_(P, z, 0x1.999999997fdb2p-2);
    P = horner_(P, z, 0x1.5555555555592p-1);
    double R = z * P;
    double lm = fma_(-s, f - R, f);
    double ed = (double)e;
    return fma_(ed, LN2_HI, fma_(ed, LN2_LO, lm));
}

// x^-0.2 for x > 0 (the step-size controller's factor, with -0.2 the double nearest -1/5), correctly
// rounded for 1e-300 <= x <= 1e300 (checked against 200-bit arithmetic; outside that range within
// 9 ulp). The controller's output feeds back into every later step, so the integrator's step
// sequence, and with it every final state, follows this function's last bit: a correctly rounded
// value agrees with glibc's pow wherever glibc's does (all but ~0.1 % of arguments).
// Method: seed y0 = exp(-0.2 log x) (~2^-49 relative), then one Newton step on x y^5 = 1 with the
// residual r = 1 - x y0^5 formed exactly by FMA two-products (double-double), y = y0 (1 + r/5)
// (error 3 r^2 / 25 ~ 2^-98), times x^-(0.2 - 1/5) = 1 - (0.2 - 1/5) ln x for the double's offset
// from -1/5 (0.2 - 1/5 = 2^-56 * 4/5 exactly).
BHM_HD double pow_m02(double x) {
    const double lx = log(x);
    const double y0 = exp(-0.2 * lx);
    if (!(x >= 1e-300 && x <= 1e300)) return y0;
    const double p2 = y0 * y0, e2 = __builtin_fma(y0, y0, -p2);
    const double p4 = p2 * p2, e4 = __builtin_fma(p2, p2, -p4) + (2.0 * p2) * e2;
    const double p5 = p4 * y0, e5 = __builtin_fma(p4, y0, -p5) + e4 * y0;
    const double p6 = p5 * x, e6 = __builtin_fma(p5, x, -p6) + e5 * x;
    const double r = (1.0 - p6) - e6;
    constexpr double DLT = 0x1.999999999999ap-57;   // 0.2 (double) - 1/5 = 2^-56 * 4/5, to 53 bits
    return y0 + y0 * (r * 0.2 - DLT * lx);
}

BHM_HD double cbrt(double x) {
    if (x == 0.0 || isnan_(x) || isinf_(x)) return x;
    double sgn = x < 0.0 ? -1.0 : 1.0;
    double ax = fabs_(x);
    int adj = 0;
    if (ax < 0x1p-1022) {
        ax = ax * 0x1p54;
        adj = -18;
    }
    uint64_t b = dbits(ax);
    int e = (int)(b >> 52) - 1023;
    double m = dfrom((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);   // [1, 2)
    int q = e >= 0 ? e / 3 : -((-e + 2) / 3);
    int rr = e - 3 * q;   // 0, 1, 2
    double w = -0x1.8b825efdf6842p-5;
    w = w * m + 0x1.3cab785e5a648p-2;
    w = w * m + -0x1.977b7db5d7b17p-1;
    w = w * m + 0x1.88cc68a4a6346p+0;
    for (int i = 0; i < 2; ++i) w = w * fma_(-m, w * w * w, 4.0) * (1.0 / 3.0);   // 5e-4 -> 6e-13
    double y = m * (w * w);
    double t = m;
    if (rr == 1) {
        y = y * CBRT2;
        t = m * 2.0;
    } else if (rr == 2) {
        y = y * CBRT4;
        t = m * 4.0;
    }
    y = fma_(t / (y * y) - y, 1.0 / 3.0, y);   // one Newton polish of y^3 = t
    int k = q + adj;
    return sgn * (y * pow2i(k));
}

// sin and cos of x (finite |x| < ~1e6 keeps full accuracy; larger |x| stays deterministic)
BHM_HD void sincos(double x, double* sp, double* cp) {
    if (isnan_(x) || isinf_(x)) {
        *sp = __builtin_nan("");
        *cp = __builtin_nan("");
        return;
    }
    double kd = floor_(fma_(x, TWO_OVER_PI, 0.5));
    double r = fma_(-kd, PIO2_3, fma_(-kd, PIO2_2, fma_(-kd, PIO2_1, x)));
    int q = (int)(kd - 4.0 * floor_(kd * 0.25));
    double z = r * r;
    // sin r = r + r z S(z) (S of degree 5), cos r = 1 + z C(z) (C of degree 6); fit errors
    // 9.0e-18 relative to r and 1.6e-18 absolute on |r| <= pi/4
    double ps = 0x1.5d6496562cae0p-33;
    ps = horner_(ps, z, -0x1.ae5db3c31d54fp-26);
    ps = horner_(ps, z, 0x1.71de34747192dp-19);
    ps = horner_(ps, z, -0x1.a01a019ab3813p-13);
    ps = horner_(ps, z, 0x1.111111110eb14p-7);
    ps = horner_(ps, z, -0x1.555555555553dp-3);
    double s = fma_(r * z, ps, r);
    double pc = -0x1.8f803897f4c42p-37;
    pc = horner_(pc, z, 0x1.1ee9844185923p-29);
    pc = horner_(pc, z, -0x1.27e4f752d538ep-22);
    pc = horner_(pc, z, 0x1.a01a019b6e90ep-16);
    pc = horner_(pc, z, -0x1.6c16c16c13f0cp-10);
    pc = horner_(pc, z, 0x1.555555555553cp-5);
    pc = fma_(pc, z, -0.5);
    double c = fma_(z, pc, 1.0);
    switch (q) {
        case 0: *sp = s; *cp = c; break;
        case 1: *sp = c; *cp = -s; break;
        case 2: *sp = -s; *cp = -c; break;
        default: *sp = -c; *cp = s; break;
    }
}

// atan(t) for t >= 0 (t may be +inf)
BHM_HD double atan_pos(double t) {
    bool inv = t > 1.0;
    if (inv) t = 1.0 / t;
    bool red = t > TAN_PI_12;
    if (red) t = (t * SQRT3 - 1.0) / (t + SQRT3);
    double z = t * t;
    // atan t = t + t z A(z), A of degree 8 (fit error 1.6e-18 relative to t)
    double p = -0x1.3ae4ad1dc9d8ap-5;
    p = horner_(p, z, 0x1.d15d3acf0adbfp-5);
    p = horner_(p, z, -0x1.1065d886df6fdp-4);
    p = horner_(p, z, 0x1.3b0b3bb362accp-4);
    p = horner_(p, z, -0x1.745cd4f68b712p-4);
    p = horner_(p, z, 0x1.c71c708c231dbp-4);
    p = horner_(p, z, -0x1.24924922ea4b0p-3);
    p = horner_(p, z, 0x1.9999999997695p-3);
    p = horner_(p, z, -0x1.555555555554dp-2);
    double a = fma_(t * z, p, t);
    if (red) a = PI_6 + a;
    if (inv) a = PI_2 - a;
    return a;
}

BHM_HD double atan2(double y, double x) {
    if (isnan_(x) || isnan_(y)) return x + y;
    double ax = fabs_(x), ay = fabs_(y);
    double a;
    if (ay == 0.0 && ax == 0.0) a = 0.0;
    else if (ay <= ax) a = atan_pos(ay / ax);
    else a = PI_2 - atan_pos(ax / ay);
    if ((dbits(x) >> 63) != 0) a = PI_ - a;
    return (dbits(y) >> 63) != 0 ? -a : a;
}

BHM_HD double acos(double x) {
    if (isnan_(x) || x > 1.0 || x < -1.0) return __builtin_nan("");
    return atan2(__builtin_sqrt((1.0 - x) * (1.0 + x)), x);
}

BHM_HD double asin(double x) {
    if (isnan_(x) || x > 1.0 || x < -1.0) return __builtin_nan("");
    return atan2(x, __builtin_sqrt((1.0 - x) * (1.0 + x)));
}

}  // namespace bhm
