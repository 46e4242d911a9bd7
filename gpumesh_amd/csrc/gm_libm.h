// gm_libm.h -- the project's own double-precision transcendentals.
//
// The smoothing path needs pow(x, 2/3) (mean-ratio corner quality,
// MeanRatioEvaluator.cpp:31), pow(x, 1/3) (GETMe volume ratio,
// GetmeSmoother.cpp:101,329) and exp (the spherical-shell metric field).
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

#endif // GM_LIBM_H
