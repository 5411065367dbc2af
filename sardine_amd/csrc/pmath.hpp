// pmath.hpp — a portable double-precision libm for the geodesy and the
// transverse Mercator projection: sin, cos, sincos, exp, log, sinh, cosh,
// atanh, atan and atan2 built only from IEEE-754 additions, subtractions,
// multiplications, divisions, square roots, explicit fused multiply-adds
// (fma) and exact scalings (rint, frexp, ldexp).  Every such operation is
// correctly rounded on the host (x86-64, fma from the C library) and on
// gfx950, so a compilation without implicit contraction (-ffp-contract=off)
// returns the same bits on both: the GPU
// kernels and their host-side planning compute the same ECEF targets and
// projected coordinates bit for bit.  Accuracy: within about one ulp (the
// tests measure each function against 50-digit values).
//
// Methods (no library code is reproduced):
//   - argument reduction by the nearest multiple of π/2 or ln 2 with the
//     constant split into two or three parts whose leading parts have 33
//     significant bits, so that k·C is exact for |k| < 2^20 (Cody and Waite,
//     "Software Manual for the Elementary Functions", 1980);
//   - Taylor polynomials on the reduced interval, evaluated by Horner's rule
//     (separate multiplies and adds: with fused ones the terrain-correction
//     chain kernel needed more registers than it has and spilled), the
//     leading term added last by a fused multiply-add; cos as
//     w + ((1 - w) - r²/2) + r⁴·C(r²), with r²/2 exact as a sum of two
//     doubles, which recovers the rounding error of w = 1 - r²/2 exactly;
//   - log by m·2^e with m in [√½, √2) and log m = 2 atanh((m-1)/(m+1));
//   - atan by a table of atan(j/8) in double-double, computed with mpmath,
//     and atan(x) = atan(c) + atan((x - c)/(1 + x c)), c = j/8 nearest x.
// Arguments are expected in the ranges the path uses (|x| below ~1e5 for
// sin/cos, below ~700 for exp); results elsewhere are not specified.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PM_HD __host__ __device__ __forceinline__
#else
#define PM_HD inline
#endif

namespace pm {

// π/2 = PIO2_1 + PIO2_2 + PIO2_3 (33 + 33 + 53 significant bits)
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
// ln 2 = LN2_1 + LN2_2
constexpr double LN2_1 = 0x1.62e42ff000000p-1;
constexpr double LN2_2 = -0x1.718432a1b0e26p-35;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
// π/2 and π as double-double
constexpr double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
constexpr double PI_HI = 0x1.921fb54442d18p+1, PI_LO = 0x1.1a62633145c07p-53;

// sin r = r + r³ S(r²) and cos r = 1 - r²/2 + r⁴ C(r²) on |r| <= π/4:
// Taylor coefficients (-1)^k/(2k+1)! for k = 1..9 and (-1)^k/(2k)! for
// k = 2..10 (truncation below 1e-19 relative).
PM_HD double sin_poly(double r) {
    const double z = r * r;
    double s = -1.0 / 121645100408832000.0;  // -1/19!
    s = s * z + 1.0 / 355687428096000.0;     //  1/17!
    s = s * z - 1.0 / 1307674368000.0;       // -1/15!
    s = s * z + 1.0 / 6227020800.0;          //  1/13!
    s = s * z - 1.0 / 39916800.0;            // -1/11!
    s = s * z + 1.0 / 362880.0;              //  1/9!
    s = s * z - 1.0 / 5040.0;                // -1/7!
    s = s * z + 1.0 / 120.0;                 //  1/5!
    s = s * z - 1.0 / 6.0;                   // -1/3!
    return fma(r, z * s, r);
}

PM_HD double cos_poly(double r) {
    const double z = r * r;
    double c = 1.0 / 2432902008176640000.0;  //  1/20!
    c = c * z - 1.0 / 6402373705728000.0;    // -1/18!
    c = c * z + 1.0 / 20922789888000.0;      //  1/16!
    c = c * z - 1.0 / 87178291200.0;         // -1/14!
    c = c * z + 1.0 / 479001600.0;           //  1/12!
    c = c * z - 1.0 / 3628800.0;             // -1/10!
    c = c * z + 1.0 / 40320.0;               //  1/8!
    c = c * z - 1.0 / 720.0;                 // -1/6!
    c = c * z + 1.0 / 24.0;                  //  1/4!
    const double hz = 0.5 * z, hz_lo = 0.5 * fma(r, r, -z);  // r²/2 = hz + hz_lo exactly
    const double w = 1.0 - hz;
    return w + fma(z * z, c, ((1.0 - w) - hz) - hz_lo);
}

// sin and cos of x with one shared reduction x = k·π/2 + r.
PM_HD void sincos(double x, double* s, double* c) {
    const double k = rint(x * TWO_OVER_PI);
    const double r = fma(-k, PIO2_3, fma(-k, PIO2_2, fma(-k, PIO2_1, x)));
    const double sr = sin_poly(r), cr = cos_poly(r);
    const int q = (int)(int64_t)k & 3;
    *s = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    *c = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
}

PM_HD double sin(double x) {
    double s, c;
    sincos(x, &s, &c);
    return s;
}

PM_HD double cos(double x) {
    double s, c;
    sincos(x, &s, &c);
    return c;
}

// e^x: x = k ln 2 + r, |r| <= ln2/2, e^r by its Taylor series to r^14.
PM_HD double exp(double x) {
    const double k = rint(x * INV_LN2);
    const double r = fma(-k, LN2_2, fma(-k, LN2_1, x));
    double p = 1.0 / 87178291200.0;  // 1/14!
    p = p * r + 1.0 / 6227020800.0;  // 1/13!
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    const double e = 1.0 + fma(r, r * p, r);
    return ldexp(e, (int)k);
}

// atanh(s) = s + s³/3 + s⁵/5 + ... for |s| <= 0.1716 (to s^23).
PM_HD double atanh_series(double s) {
    const double z = s * s;
    double p = 1.0 / 23.0;
    p = p * z + 1.0 / 21.0;
    p = p * z + 1.0 / 19.0;
    p = p * z + 1.0 / 17.0;
    p = p * z + 1.0 / 15.0;
    p = p * z + 1.0 / 13.0;
    p = p * z + 1.0 / 11.0;
    p = p * z + 1.0 / 9.0;
    p = p * z + 1.0 / 7.0;
    p = p * z + 1.0 / 5.0;
    p = p * z + 1.0 / 3.0;
    return fma(s, z * p, s);
}

// log y for finite y > 0.
PM_HD double log(double y) {
    int e = 0;
    double m = frexp(y, &e);  // y = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752) {
        m = m * 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);  // m - 1 is exact
    const double ed = (double)e;
    return fma(ed, LN2_1, fma(ed, LN2_2, 2.0 * atanh_series(s)));
}

// atanh x for |x| < 1.
PM_HD double atanh(double x) {
    const double a = fabs(x);
    const double r = a <= 0.1716 ? atanh_series(a) : 0.5 * log((1.0 + a) / (1.0 - a));
    return x < 0.0 ? -r : r;
}

// sinh x: its Taylor series below 1/2 (to x^19), (e^|x| - e^-|x|)/2 above.
PM_HD double sinh(double x) {
    const double a = fabs(x);
    double r;
    if (a < 0.5) {
        const double z = a * a;
        double p = 1.0 / 121645100408832000.0;  // 1/19!
        p = p * z + 1.0 / 355687428096000.0;
        p = p * z + 1.0 / 1307674368000.0;
        p = p * z + 1.0 / 6227020800.0;
        p = p * z + 1.0 / 39916800.0;
        p = p * z + 1.0 / 362880.0;
        p = p * z + 1.0 / 5040.0;
        p = p * z + 1.0 / 120.0;
        p = p * z + 1.0 / 6.0;
        r = fma(a, z * p, a);
    } else {
        const double e = exp(a);
        r = 0.5 * (e - 1.0 / e);
    }
    return x < 0.0 ? -r : r;
}

PM_HD double cosh(double x) {
    const double e = exp(fabs(x));
    return 0.5 * (e + 1.0 / e);
}

// atan(j/8), j = 0..8, as double-double (hi + lo)
PM_HD double atan_table_hi(int j) {
    static constexpr double t[9] = {0.0, 0x1.fd5ba9aac2f6ep-4, 0x1.f5b75f92c80ddp-3, 0x1.6f61941e4def1p-2,
                             0x1.dac670561bb4fp-2, 0x1.1e00babdefeb4p-1, 0x1.4978fa3269ee1p-1,
                             0x1.700a7c5784634p-1, 0x1.921fb54442d18p-1};
    return t[j];
}
PM_HD double atan_table_lo(int j) {
    static constexpr double t[9] = {0.0, -0x1.cd37686760c17p-59, 0x1.8ab6e3cf7afbdp-57, -0x1.c63aae6f6e918p-56,
                             0x1.a2b7f222f65e2p-56, -0x1.928df287a668fp-58, 0x1.2419a87f2a458p-56,
                             -0x1.8c34d25aadef6p-56, 0x1.1a62633145c07p-55};
    return t[j];
}

// atan z - z for |z| <= 1/16 (series to z^17)
PM_HD double atan_tail(double z) {
    const double w = z * z;
    double p = 1.0 / 17.0;
    p = p * w - 1.0 / 15.0;
    p = p * w + 1.0 / 13.0;
    p = p * w - 1.0 / 11.0;
    p = p * w + 1.0 / 9.0;
    p = p * w - 1.0 / 7.0;
    p = p * w + 1.0 / 5.0;
    p = p * w - 1.0 / 3.0;
    return z * (w * p);
}

PM_HD double atan(double x) {
    double a = fabs(x);
    const bool inv = a > 1.0;
    if (inv) a = 1.0 / a;
    const int j = (int)rint(a * 8.0);
    const double c = (double)j * 0.125;
    const double z = (a - c) / fma(a, c, 1.0);  // a - c is exact
    double r = atan_table_hi(j) + (atan_table_lo(j) + (z + atan_tail(z)));
    if (inv) r = (PIO2_HI - r) + PIO2_LO;
    return x < 0.0 ? -r : r;
}

PM_HD double atan2(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    if (ax == 0.0 && ay == 0.0) return 0.0;
    double r = ay <= ax ? atan(ay / ax) : (PIO2_HI - atan(ax / ay)) + PIO2_LO;
    if (x < 0.0) r = (PI_HI - r) + PI_LO;
    return y < 0.0 ? -r : r;
}

}  // namespace pm
