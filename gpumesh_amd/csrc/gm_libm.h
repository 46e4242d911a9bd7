// gm_libm.h -- the project's own double-precision transcendentals.
//
// The smoothing path needs pow(x, 2/3) (mean-ratio corner quality,
// MeanRatioEvaluator.cpp:31), pow(x, 1/3) (GETMe volume ratio,
// GetmeSmoother.cpp:101,329), exp (the spherical-shell metric field), and
// cos and pow(x, y) (the reference's analytic field,
// Samplers/AbstractSampler.cpp:190-208).
// glibc and the device math library may disagree in the last ulp, and a
// last-ulp difference inside an optimiser's comparison sends a vertex down a
// different branch.  Every routine here is built only from IEEE-754 +, -, *,
// / and exponent bit manipulation, so it returns the same bits on an x86-64
// host and on a gfx950 wavefront as long as both sides are compiled without
// floating-point contraction (-ffp-contract=off) and without fast-math.
//
// Methods (all textbook, none taken from a vendor library):
//   * cube root: exponent split x = t * 2^(3q), t in [1, 8); cubic
//     least-squares start (max rel. error 1.4e-2); two Halley steps
//     y <- y (y^3 + 2t) / (2y^3 + t) (cubic convergence); one Newton step on
//     the residual t - y^3 formed exactly with Dekker's product.
//   * exp: Cody-Waite reduction x = n ln2 + r with a two-part ln2, |r| <= ln2/2,
//     degree-13 Taylor polynomial in Horner form, exact scaling by 2^n.
//   * cos: Cody-Waite reduction by pi/2 (three parts), Taylor series of
//     cos r / sin r by quadrant.
//   * ln: exponent split, ln m = 2 atanh((m-1)/(m+1)) series; pow(x, y) =
//     exp(y ln x).
// Accuracy against extended-precision values is measured by the unit tests
// (within 1 ulp).
#ifndef GM_LIBM_H
#define GM_LIBM_H

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define GM_HD __host__ __device__ inline
#else
#define GM_HD static inline
#endif

GM_HD uint64_t gm_dbits(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof(u));
    return u;
}

GM_HD double gm_dfrom(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof(x));
    return x;
}

// 2^k for -1022 <= k <= 1023 (exact).
GM_HD double gm_pow2i(int k)
{
    return gm_dfrom((uint64_t)(k + 1023) << 52);
}

// Splits x > 0 (finite) into t in [1, 8) and q with x = t * 2^(3q).
GM_HD double gm_cbrt_reduce(double x, int* q_out)
{
    uint64_t u = gm_dbits(x);
    int e = (int)((u >> 52) & 0x7ff);
    int adj = 0;
    if(e == 0)
    {
        // Subnormal: scale into the normal range first (exact).
        x = x * 18014398509481984.0; // 2^54
        u = gm_dbits(x);
        e = (int)((u >> 52) & 0x7ff);
        adj = -54;
    }
    int E = e - 1023 + adj;
    int q = (E >= 0) ? (E / 3) : -((-E + 2) / 3);
    int r = E - 3 * q;
    *q_out = q;
    return gm_dfrom((u & 0x000fffffffffffffULL) | ((uint64_t)(1023 + r) << 52));
}

// Exact product a*b = p + e without FMA (Dekker: Veltkamp split into 26-bit halves).
GM_HD void gm_two_prod(double a, double b, double* p, double* e)
{
    const double SPLIT = 134217729.0; // 2^27 + 1
    double ca = SPLIT * a, cb = SPLIT * b;
    double ah = ca - (ca - a), al = a - ah;
    double bh = cb - (cb - b), bl = b - bh;
    *p = a * b;
    *e = ((ah * bh - *p) + ah * bl + al * bh) + al * bl;
}

// For t in [1, 8): y ~ t^(1/3) from two Halley steps, then one Newton step
// on an exactly computed residual r = t - y^3 (error-free products), which
// leaves y + delta with delta = r / (3 y^2) accurate to well below an ulp.
// Returns y and delta separately so that callers can form y^2 accurately.
GM_HD double gm_cbrt_core(double t, double* delta)
{
    double y = 0.7091206995822162 + t * (0.3400291293011452 +
               t * (-0.03749812135376207 + t * 0.001916834543139507));
    for(int i = 0; i < 2; ++i)
    {
        double y3 = y * y * y;
        y = y * (y3 + 2.0 * t) / (2.0 * y3 + t);
    }
    double s, se, p1, e1;
    gm_two_prod(y, y, &s, &se);
    gm_two_prod(y, s, &p1, &e1);
    double r = (t - p1) - (e1 + y * se);
    *delta = r / (3.0 * s);
    return y;
}

// x^(1/3) for x >= 0; NaN for x < 0 or NaN (pow semantics); +inf -> +inf.
GM_HD double gm_cbrt_nonneg(double x)
{
    if(!(x > 0.0))
        return (x == 0.0) ? 0.0 : (x - x) / (x - x);
    if(x == (double)__builtin_huge_val())
        return x;
    int q;
    double t = gm_cbrt_reduce(x, &q);
    double d;
    double y = gm_cbrt_core(t, &d);
    return (y + d) * gm_pow2i(q);
}

// x^(2/3) for x >= 0 (the mean-ratio exponent); NaN for x < 0 or NaN.
GM_HD double gm_pow23(double x)
{
    if(!(x > 0.0))
        return (x == 0.0) ? 0.0 : (x - x) / (x - x);
    if(x == (double)__builtin_huge_val())
        return x;
    int q;
    double t = gm_cbrt_reduce(x, &q);
    double d;
    double y = gm_cbrt_core(t, &d);
    // (y + d)^2 = s + se + 2 y d + d^2 with y^2 = s + se exactly.
    double s, se;
    gm_two_prod(y, y, &s, &se);
    return (s + (se + 2.0 * y * d)) * gm_pow2i(2 * q);
}

// e^x.
GM_HD double gm_exp(double x)
{
    if(x != x)
        return x;
    if(x > 709.78)
        return (double)__builtin_huge_val();
    if(x < -745.2)
        return 0.0;
    const double LOG2E = 1.4426950408889634;
    const double LN2_HI = 6.93147180369123816490e-01; // 32 significant bits
    const double LN2_LO = 1.90821492927058770002e-10;
    double nf = x * LOG2E;
    nf = (nf >= 0.0) ? (double)(int64_t)(nf + 0.5) : -(double)(int64_t)(-nf + 0.5);
    int n = (int)nf;
    double r = (x - nf * LN2_HI) - nf * LN2_LO;
    // Taylor series of e^r, |r| <= 0.347: the degree-13 remainder is < 2e-18.
    double p = 1.0 / 6227020800.0;               // 1/13!
    p = 1.0 / 479001600.0 + r * p;               // 1/12!
    p = 1.0 / 39916800.0 + r * p;                // 1/11!
    p = 1.0 / 3628800.0 + r * p;                 // 1/10!
    p = 1.0 / 362880.0 + r * p;                  // 1/9!
    p = 1.0 / 40320.0 + r * p;                   // 1/8!
    p = 1.0 / 5040.0 + r * p;                    // 1/7!
    p = 1.0 / 720.0 + r * p;                     // 1/6!
    p = 1.0 / 120.0 + r * p;                     // 1/5!
    p = 1.0 / 24.0 + r * p;                      // 1/4!
    p = 1.0 / 6.0 + r * p;                       // 1/3!
    p = 0.5 + r * p;                             // 1/2!
    p = 1.0 + r * p;                             // 1/1!
    p = 1.0 + r * p;                             // 1/0!
    if(n > 1023)
        return p * gm_pow2i(1023) * gm_pow2i(n - 1023);
    if(n < -1022)
        return p * gm_pow2i(-1022) * gm_pow2i(n + 1022);
    return p * gm_pow2i(n);
}

// Rounds x to the nearest integer (halves away from zero), |x| < 2^62.
GM_HD double gm_round_half_away(double x)
{
    return (x >= 0.0) ? (double)(int64_t)(x + 0.5) : -(double)(int64_t)(-x + 0.5);
}

// cos(x): Cody-Waite reduction x = n pi/2 + r, with pi/2 split into parts of
// 33, 33 and 53 significant bits (n pi/2 is formed exactly while |n| < 2^20,
// so the reduction keeps full accuracy for |x| < 2^19 pi; beyond that it
// loses bits, identically on every platform), then the Taylor series of
// cos r or sin r for |r| <= pi/4 (remainders below 4e-21) chosen by the
// quadrant n mod 4.  NaN for infinite or NaN input.
GM_HD double gm_cos(double x)
{
    if(x != x || x - x != 0.0)
        return (x - x) / (x - x);
    const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
    const double PIO2_1 = 0x1.921fb54400000p+0;
    const double PIO2_2 = 0x1.0b4611a600000p-34;
    const double PIO2_3 = 0x1.3198a2e037073p-69;
    const double nf = gm_round_half_away(x * TWO_OVER_PI);
    const double r = ((x - nf * PIO2_1) - nf * PIO2_2) - nf * PIO2_3;
    const double z = r * r;
    const int q = (int)((int64_t)nf & 3);
    double t;
    if((q & 1) == 0)
    {
        // cos r = sum (-1)^k z^k / (2k)!
        t = -1.0 / 6402373705728000.0;        // 18!
        t = 1.0 / 20922789888000.0 + z * t;    // 16!
        t = -1.0 / 87178291200.0 + z * t;      // 14!
        t = 1.0 / 479001600.0 + z * t;         // 12!
        t = -1.0 / 3628800.0 + z * t;          // 10!
        t = 1.0 / 40320.0 + z * t;             // 8!
        t = -1.0 / 720.0 + z * t;              // 6!
        t = 1.0 / 24.0 + z * t;                // 4!
        t = -0.5 + z * t;                      // 2!
        t = 1.0 + z * t;
    }
    else
    {
        // sin r = r sum (-1)^k z^k / (2k+1)!
        t = -1.0 / 121645100408832000.0;       // 19!
        t = 1.0 / 355687428096000.0 + z * t;   // 17!
        t = -1.0 / 1307674368000.0 + z * t;    // 15!
        t = 1.0 / 6227020800.0 + z * t;        // 13!
        t = -1.0 / 39916800.0 + z * t;         // 11!
        t = 1.0 / 362880.0 + z * t;            // 9!
        t = -1.0 / 5040.0 + z * t;             // 7!
        t = 1.0 / 120.0 + z * t;               // 5!
        t = -1.0 / 6.0 + z * t;                // 3!
        t = r + r * (z * t);
    }
    // cos(n pi/2 + r) = cos r, -sin r, -cos r, sin r for n mod 4 = 0..3.
    return (q == 0 || q == 3) ? t : -t;
}

// ln x: x = 2^k m with m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with
// s = (m - 1)/(m + 1), |s| <= 0.1716 (the series in s^2 to degree 12, the
// remainder below 1e-21), k ln 2 from a two-part ln 2 (42 + 53 bits, k ln2_hi
// exact).  ln 0 = -inf, ln of a negative number or NaN = NaN, ln inf = inf.
GM_HD double gm_log(double x)
{
    if(!(x > 0.0))
        return (x == 0.0) ? -(double)__builtin_huge_val() : (x - x) / (x - x);
    if(x == (double)__builtin_huge_val())
        return x;
    uint64_t u = gm_dbits(x);
    int e = (int)((u >> 52) & 0x7ff);
    int adj = 0;
    if(e == 0)
    {
        x = x * 18014398509481984.0;   // 2^54, exact
        u = gm_dbits(x);
        e = (int)((u >> 52) & 0x7ff);
        adj = -54;
    }
    int k = e - 1023 + adj;
    double m = gm_dfrom((u & 0x000fffffffffffffULL) | ((uint64_t)1023 << 52));   // [1, 2)
    if(m > 0x1.6a09e667f3bcdp+0)
    {
        m = 0.5 * m;
        ++k;
    }
    const double f = m - 1.0;        // exact (m within a factor 2 of 1)
    const double s = f / (2.0 + f);
    const double z = s * s;
    double t = 1.0 / 25.0;
    t = 1.0 / 23.0 + z * t;
    t = 1.0 / 21.0 + z * t;
    t = 1.0 / 19.0 + z * t;
    t = 1.0 / 17.0 + z * t;
    t = 1.0 / 15.0 + z * t;
    t = 1.0 / 13.0 + z * t;
    t = 1.0 / 11.0 + z * t;
    t = 1.0 / 9.0 + z * t;
    t = 1.0 / 7.0 + z * t;
    t = 1.0 / 5.0 + z * t;
    t = 1.0 / 3.0 + z * t;
    const double lm = 2.0 * s + 2.0 * s * (z * t);
    const double LN2_HI = 0x1.62e42fefa3800p-1;
    const double LN2_LO = 0x1.ef35793c76730p-45;
    const double kd = (double)k;
    return kd * LN2_HI + (kd * LN2_LO + lm);
}

// x^y for x >= 0 as exp(y ln x): x = 0 gives 0 (y > 0), 1 (y = 0) or inf
// (y < 0); x = 1 or y = 0 gives 1; a negative x gives NaN.
GM_HD double gm_pow(double x, double y)
{
    if(y == 0.0 || x == 1.0)
        return 1.0;
    if(x == 0.0)
        return y > 0.0 ? 0.0 : (double)__builtin_huge_val();
    if(!(x > 0.0) || y != y)
        return (x - x) / (x - x) + y;
    return gm_exp(y * gm_log(x));
}

#endif // GM_LIBM_H
