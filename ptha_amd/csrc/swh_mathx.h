// swh_mathx.h -- FP64 pow and sin accurate to about 2^-70 relative before the final
// rounding, for the places where the reference calls libm: Manning/Chezy friction
// depth^(-7/3) (domain_mod_update_U_DE1_inner_include.f90:74-75, the leapfrog friction
// work, domain_mod_leapfrog_nonlinear_include.f90:882) and the stage-function sine of
// the boundary conditions (boundary_mod.f90:78-164).
//
// The reference's libm (glibc 2.35) rounds pow and sin almost always correctly; the
// closer a device result is to the correctly rounded one, the fewer ulp differences
// seed round-off divergence over long wet/dry runs. These functions evaluate the exact
// value to ~2^-70 relative and round once, so they return the correctly rounded result
// except where the exact value lies within ~2^-70 of a rounding midpoint.
//
// Methods (published):
//  * log: Tang's table-driven logarithm (P. T. P. Tang, ACM TOMS 16(4), 1990):
//    x = 2^e m, m in [1, 2), F = 1 + j/128 nearest m, f = m - F (exact),
//    log x = e ln2 + log F + 2 atanh(f / (2F + f)), evaluated in double-double.
//  * exp: Tang's table-driven exponential (ACM TOMS 15(2), 1989): N = round(v 64/ln2),
//    r = v - N ln2/64 (double-double), exp v = 2^(N div 64) 2^((N mod 64)/64) e^r,
//    e^r - 1 a degree-7 Taylor polynomial (|r| <= ln2/128).
//  * pow(x, y) = exp(y log x) with y log x carried in double-double.
//  * sin: Cody-Waite reduction by pi/2 in four parts, then Taylor series of sin or cos
//    evaluated entirely in double-double.
// Only IEEE basic operations and fma are used, so the host build (used to check these
// functions against glibc and against quad-precision values on the CPU) and the device
// build give the same bits. Arguments outside the ranges handled (x <= 0, non-finite,
// results outside the normal range, |t| > 2^20) go to the library function.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SWHM_HD __host__ __device__
#else
#define SWHM_HD
#endif

#include "swh_mathx_tables.h"

namespace swhm {

struct dd {
    double hi, lo;
};

SWHM_HD inline dd two_sum(double a, double b)
{
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}

// |a| >= |b| (or a == 0)
SWHM_HD inline dd fast_two_sum(double a, double b)
{
    const double s = a + b;
    return {s, b - (s - a)};
}

SWHM_HD inline dd two_prod(double a, double b)
{
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

SWHM_HD inline dd dd_add(dd a, dd b)
{
    const dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    dd u = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(u.hi, u.lo + t.lo);
}

SWHM_HD inline dd dd_mul(dd a, dd b)
{
    dd p = two_prod(a.hi, b.hi);
    const double c = a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo + c);
}

SWHM_HD inline dd dd_mul_d(dd a, double b)
{
    dd p = two_prod(a.hi, b);
    return fast_two_sum(p.hi, p.lo + a.lo * b);
}

SWHM_HD inline int64_t bits_of(double x)
{
    int64_t b;
    memcpy(&b, &x, sizeof b);
    return b;
}

SWHM_HD inline double from_bits(int64_t b)
{
    double x;
    memcpy(&x, &b, sizeof x);
    return x;
}

// log(x) as a double-double with absolute error ~2^-78, x > 0 finite.
SWHM_HD inline dd log_dd(double x)
{
    int64_t b = bits_of(x);
    int e = 0;
    if ((b >> 52) == 0) {   // subnormal: scale into the normal range
        x = x * 0x1p54;
        b = bits_of(x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    const double m = from_bits((b & 0x000fffffffffffffLL) | 0x3ff0000000000000LL);   // [1, 2)
    const int j = (int)((m - 1.0) * 128.0 + 0.5);                                       // 0..128
    const double F = 1.0 + (double)j * 0x1p-7;
    const double f = m - F;   // exact: |f| <= 2^-8, both multiples of 2^-52
    // s = f / (2F + f) as a double-double
    const double num = f;
    const dd den = two_sum(2.0 * F, f);
    const double sh = num / den.hi;
    const double sl = (fma(-sh, den.hi, num) - sh * den.lo) / den.hi;
    // 2 atanh(s) = 2s + (2/3) s^3 + (2/5) s^5 + (2/7) s^7 + ..., |s| <= 2^-9
    const double q = sh * sh;
    const double tail = sh * q * (0x1.5555555555555p-1 + q * (0x1.999999999999ap-2 + q * 0x1.2492492492492p-2));
    const double ed = (double)e;
    const dd a = two_sum(ed * kLn2Hi, kLogF[j][0]);
    const dd c = two_sum(a.hi, 2.0 * sh);
    const double lo = a.lo + c.lo + (ed * kLn2Lo + kLogF[j][1] + 2.0 * sl + tail);
    return fast_two_sum(c.hi, lo);
}

// exp(v.hi + v.lo) rounded once to double; v.hi within [-700, 700].
SWHM_HD inline double exp_dd_round(dd v)
{
    const double Nd = rint(v.hi * kInvLn2o64);
    const int N = (int)Nd;
    const int j = N & 63;
    const int M = (N - j) / 64;
    // r = v - N ln2/64 as a double-double (N * kLn2o64Hi is exact)
    const dd t = two_sum(v.hi, -Nd * kLn2o64Hi);
    const dd r = fast_two_sum(t.hi, t.lo + v.lo - Nd * kLn2o64Lo);
    // e^r - 1 = r + r^2/2 + r^3 (1/6 + r/24 + ... + r^4/5040), |r| <= ln2/128
    dd r2 = two_prod(r.hi, r.hi);
    r2.lo += 2.0 * r.hi * r.lo;
    r2 = fast_two_sum(0.5 * r2.hi, 0.5 * r2.lo);
    const double rh = r.hi;
    const double q = rh * rh * rh *
        (0x1.5555555555555p-3 +
         rh * (0x1.5555555555555p-5 + rh * (0x1.1111111111111p-7 + rh * (0x1.6c16c16c16c17p-10 + rh * 0x1.a01a01a01a01ap-13))));
    dd p = dd_add(r, r2);
    p = fast_two_sum(p.hi, p.lo + q);
    // 2^(j/64) (1 + p)
    const double Th = kExp2J[j][0], Tl = kExp2J[j][1];
    const dd x1 = two_prod(Th, p.hi);
    const dd s = two_sum(Th, x1.hi);
    const double lo = s.lo + x1.lo + Tl + (Th * p.lo + Tl * p.hi);
    const double y = s.hi + lo;   // the one rounding
    return ldexp(y, M);
}

// pow(x, y) for x > 0 (used with y = -7/3 and -2).
SWHM_HD inline double pow_acc(double x, double y)
{
    if (!(x > 0.0) || !(x < INFINITY) || !(fabs(y) < INFINITY)) return pow(x, y);
    const dd l = log_dd(x);
    const dd v = dd_mul_d(l, y);
    if (!(fabs(v.hi) <= 700.0)) return pow(x, y);
    return exp_dd_round(v);
}

// sin(r) or cos(r) for a double-double |r| <= pi/4 (+ tiny), as a double-double
// (Taylor series with every term in double-double, to r^27 / 27!).
SWHM_HD inline dd sincos_dd(dd r, bool want_cos)
{
    const dd r2 = dd_mul(r, r);
    // Horner from the top: sin r = r (1 - r^2/(2*3) (1 - r^2/(4*5) (1 - ...)))
    //                      cos r = 1 - r^2/(1*2) (1 - r^2/(3*4) (1 - ...))
    dd acc = {1.0, 0.0};
    const int top = want_cos ? 26 : 27;
    for (int k = top; k >= 2; k -= 2) {
        // acc = 1 - r^2 / (k (k-1)) * acc
        const double den = (double)k * (double)(k - 1);
        dd t = dd_mul(r2, acc);
        // t / den as a double-double
        const double qh = t.hi / den;
        const double ql = (fma(-qh, den, t.hi) + t.lo) / den;
        acc = dd_add({1.0, 0.0}, {-qh, -ql});
    }
    return want_cos ? acc : dd_mul(r, acc);
}

// sin(t) rounded once to double.
SWHM_HD inline double sin_acc(double t)
{
    if (!(fabs(t) <= 0x1p20)) return sin(t);
    const double kd = rint(t * kTwoOverPi);
    // r = t - k pi/2 with pi/2 in four parts (k * part exact for |k| < 2^20)
    dd r = two_sum(t, -kd * kPio2_1);
    r = dd_add(r, {-kd * kPio2_2, 0.0});
    r = dd_add(r, {-kd * kPio2_3, 0.0});
    r = dd_add(r, {-kd * kPio2_4, 0.0});
    const int q = ((int)kd) & 3;
    dd v = sincos_dd(r, (q & 1) != 0);
    if (q >= 2) v = {-v.hi, -v.lo};
    return v.hi + v.lo;
}

}  // namespace swhm
