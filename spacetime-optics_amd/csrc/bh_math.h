/* bh_math.h - FP64 elementary functions that give bit-identical results on the x86-64 host
   (g++, SSE2) and on gfx950 (hipcc), provided both sides compile with -ffp-contract=off.

   Why: the adaptive integrator branches on values that pass through pow/exp/log/atan2 every step
   (step-size controller, swept angle, emission). Near the photon ring the geodesic flow is
   unstable (Lyapunov factor ~e^pi per half orbit), so a one-ulp difference between two libm
   implementations grows into different step counts, winding counts and even fates. Routing every
   per-step transcendental through this header makes the GPU tracer and the CPU oracle agree bit
   for bit; accuracy against glibc is <= 3 ulp (pinned by tests/test_bh_math.py over 10^6 inputs per
   function), and x^-0.2 is correctly rounded.

   Methods (all textbook, coefficients are exact series terms rounded to double):
   - exp, expm1: Cody-Waite reduction x = k ln2 + r with a 32-bit ln2 head, |r| <= ln2/2, Taylor
     series of degree 13 (exp) / 14 (expm1) in Horner form; expm1 recombines 2^k expm1(r) + (2^k-1).
   - log: x = m 2^e, m in [sqrt(1/2), sqrt(2)], f = m-1, s = f/(2+f), log(1+f) = f - s (f - R) with
     R = sum_{n>=1} 2 s^{2n}/(2n+1) (the atanh series), to n = 12.
   - cbrt: m in [1,2) and the exponent split by 3; Newton iteration w <- w (4 - m w^3)/3 for
     m^{-1/3} from a cubic least-squares seed, then cbrt(m) = m w^2.
   - sincos: Cody-Waite reduction by pi/2 with a three-part constant (33+33+53 bits), Taylor series
     on |r| <= pi/4 (sin to r^17, cos to r^18).
   - atan2/acos: reciprocal and pi/6 (Hart) reductions to |t| <= tan(pi/12), odd Taylor series to
     t^29; acos(x) = atan2(sqrt((1-x)(1+x)), x), asin(x) = atan2(x, sqrt((1-x)(1+x))).
   - x^-0.2 (the step controller): exp(-0.2 log x) refined by one Newton step on x y^5 = 1 with an
     FMA double-double residual; correctly rounded on [1e-300, 1e300]. FMA is exact on both sides
     (v_fma_f64 on the device, the C library's fma on the host), so this stays bit-identical.
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

// 1/n!
constexpr double F2 = 0x1.0000000000000p-1, F3 = 0x1.5555555555555p-3, F4 = 0x1.5555555555555p-5,
                 F5 = 0x1.1111111111111p-7, F6 = 0x1.6c16c16c16c17p-10, F7 = 0x1.a01a01a01a01ap-13,
                 F8 = 0x1.a01a01a01a01ap-16, F9 = 0x1.71de3a556c734p-19, F10 = 0x1.27e4fb7789f5cp-22,
                 F11 = 0x1.ae64567f544e4p-26, F12 = 0x1.1eed8eff8d898p-29, F13 = 0x1.6124613a86d09p-33,
                 F14 = 0x1.93974a8c07c9dp-37, F15 = 0x1.ae7f3e733b81fp-41, F16 = 0x1.ae7f3e733b81fp-45,
                 F17 = 0x1.952c77030ad4ap-49, F18 = 0x1.6827863b97d97p-53;

// e^r - 1 - r for |r| <= ln2/2, as r^2 * q(r)
BHM_HD double expm1_core_q(double r) {
    double p = F14;
    p = p * r + F13;
    p = p * r + F12;
    p = p * r + F11;
    p = p * r + F10;
    p = p * r + F9;
    p = p * r + F8;
    p = p * r + F7;
    p = p * r + F6;
    p = p * r + F5;
    p = p * r + F4;
    p = p * r + F3;
    p = p * r + F2;
    return p;
}

BHM_HD double exp(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -708.3964185322641) return 0.0;   // result would be subnormal: flushed
    double kd = floor_(x * INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double e = 1.0 + (r + (r * r) * expm1_core_q(r));
    if (k > 1023) return e * 2.0 * pow2i(k - 1);
    if (k < -1021) return (e * pow2i(k + 54)) * 0x1p-54;
    return e * pow2i(k);
}

BHM_HD double expm1(double x) {
    if (isnan_(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -38.0) return -1.0;   // e^x < 2^-54
    if (fabs_(x) <= 0.34657359027997264) return x + (x * x) * expm1_core_q(x);
    double kd = floor_(x * INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double em = r + (r * r) * expm1_core_q(r);
    if (k > 56) {
        double e = em + 1.0;
        if (k > 1023) return e * 2.0 * pow2i(k - 1);
        return e * pow2i(k);
    }
    double t = pow2i(k);
    return t * em + (t - 1.0);
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
    double R = 0x1.47ae147ae147bp-4;   // 2/25
    R = R * z + 0x1.642c8590b2164p-4;  // 2/23
    R = R * z + 0x1.8618618618618p-4;  // 2/21
    R = R * z + 0x1.af286bca1af28p-4;  // 2/19
    R = R * z + 0x1.e1e1e1e1e1e1ep-4;  // 2/17
    R = R * z + 0x1.1111111111111p-3;  // 2/15
    R = R * z + 0x1.3b13b13b13b14p-3;  // 2/13
    R = R * z + 0x1.745d1745d1746p-3;  // 2/11
    R = R * z + 0x1.c71c71c71c71cp-3;  // 2/9
    R = R * z + 0x1.2492492492492p-2;  // 2/7
    R = R * z + 0x1.999999999999ap-2;  // 2/5
    R = R * z + 0x1.5555555555555p-1;  // 2/3
    R = R * z;
    double lm = f - s * (f - R);
    double ed = (double)e;
    return ed * LN2_HI + (lm + ed * LN2_LO);
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
    for (int i = 0; i < 3; ++i) w = w * (4.0 - m * (w * w * w)) * (1.0 / 3.0);
    double y = m * (w * w);
    double t = m;
    if (rr == 1) {
        y = y * CBRT2;
        t = m * 2.0;
    } else if (rr == 2) {
        y = y * CBRT4;
        t = m * 4.0;
    }
    y = y + (t / (y * y) - y) * (1.0 / 3.0);   // one Newton polish of y^3 = t
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
    double kd = floor_(x * TWO_OVER_PI + 0.5);
    double r = ((x - kd * PIO2_1) - kd * PIO2_2) - kd * PIO2_3;
    int q = (int)(kd - 4.0 * floor_(kd * 0.25));
    double z = r * r;
    // sin r = r + r z (-1/3! + z (1/5! - ... + z/17!)), cos r = 1 + z (-1/2 + z (1/4! - ... - z/18!))
    double ps = F17;
    ps = ps * z - F15;
    ps = ps * z + F13;
    ps = ps * z - F11;
    ps = ps * z + F9;
    ps = ps * z - F7;
    ps = ps * z + F5;
    ps = ps * z - F3;
    double s = r + (r * z) * ps;
    double pc = -F18;
    pc = pc * z + F16;
    pc = pc * z - F14;
    pc = pc * z + F12;
    pc = pc * z - F10;
    pc = pc * z + F8;
    pc = pc * z - F6;
    pc = pc * z + F4;
    double c = 1.0 + z * (pc * z - F2);
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
    double p = 0x1.1a7b9611a7b96p-5;     // 1/29
    p = p * z - 0x1.2f684bda12f68p-5;    // 1/27
    p = p * z + 0x1.47ae147ae147bp-5;    // 1/25
    p = p * z - 0x1.642c8590b2164p-5;    // 1/23
    p = p * z + 0x1.8618618618618p-5;    // 1/21
    p = p * z - 0x1.af286bca1af28p-5;    // 1/19
    p = p * z + 0x1.e1e1e1e1e1e1ep-5;    // 1/17
    p = p * z - 0x1.1111111111111p-4;    // 1/15
    p = p * z + 0x1.3b13b13b13b14p-4;    // 1/13
    p = p * z - 0x1.745d1745d1746p-4;    // 1/11
    p = p * z + 0x1.c71c71c71c71cp-4;    // 1/9
    p = p * z - 0x1.2492492492492p-3;    // 1/7
    p = p * z + 0x1.999999999999ap-3;    // 1/5
    p = p * z - 0x1.5555555555555p-2;    // 1/3
    double a = t + (t * z) * p;
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
