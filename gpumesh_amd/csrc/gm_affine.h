// gm_affine.h -- patch evaluation of the mean-ratio evaluator with the
// metric-free measurer for optimisers that try many positions of one vertex
// while the rest of its patch stays put: the FP32 passes and the FP64 polish
// pass of the mixed-precision mode.
//
// Every mean-ratio corner of an element (the tetrahedron's single corner,
// MeanRatioEvaluator::tetQuality :34-50, or a hexahedron corner, :106-167) is
// q = sign(d) * 3 |d|^(2/3) / f2 with F the corner's 3x3 edge matrix,
// d = det F and f2 = |F|_F^2.  When only the vertex being relocated moves, by
// x from its current position, every column of F changes by r_j x with
// r_j in {-1, 0, 1} (times the constant Fr_TET_INV for tetrahedra), so
//   F(x) = F0 + x r^T,
//   det F(x) = d0 + g . x            (the rank-one determinant identity),
//   |F(x)|^2 = c0 + h . x + s |x|^2,
// with d0 = det F0, g = sum_j r_j (c_{j+1} x c_{j+2}), c0 = |F0|^2,
// h = 2 sum_j r_j c_j, s = |r|^2, c_j the columns of F0.  Staging these nine
// numbers per corner once, with coordinates taken relative to the vertex's
// position, turns a corner evaluation into a handful of FMAs and one
// |d|^(-2/3).  The identity is exact; results differ from the direct formula
// only by rounding.  Mean ratio is invariant under uniform scaling, so the
// metric scaling factor drops out.
//
// Patch quality (AbstractEvaluator::patchQuality, accumulatePatchQuality and
// finalizePatchQuality, AbstractEvaluator.cpp:778-803) is the harmonic mean of
// the element qualities if all are positive, else the smallest non-positive
// one; 1/q = f2 |d|^(-2/3) / 3 needs no division.
#ifndef GM_AFFINE_H
#define GM_AFFINE_H

#include "gm_quality.h"

namespace gma
{

using gmq::v3;
using gmq::mk;

// A corner whose FP32 determinant is below thr = NEAR_DEGENERATE * c0^(3/2)
// in magnitude (quality below about 0.006) is near-degenerate: its sign --
// valid or inverted -- is decided in FP64 from the FP32 positions (for the
// constant corners of a staged hexahedron, and for every corner at a
// vertex's chosen new position).  FP32 rounding of d is about 1e-7 c0^(3/2),
// far inside the band.
constexpr float NEAR_DEGENERATE = 1.0e-4f;

// One corner: d0, g (3), c0, h (3), s, then (FP32 only) the near-degenerate
// threshold and the corner's identity (element reference, hex corner index)
// for the FP64 determinant recheck.  48 bytes in FP32, 96 in FP64: whole
// 16-byte rows for vector LDS loads.
template<typename T>
struct alignas(16) Corner
{
    T d0, gx, gy, gz, c0, hx, hy, hz, s;
    float thr;
    uint32_t ref;
    uint32_t corner;
};

#define GMA __host__ __device__ inline

#if defined(__HIP_DEVICE_COMPILE__)
GMA float flog2(float x) { return __builtin_amdgcn_logf(x); }
GMA float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
GMA float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
GMA float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
GMA float flog2(float x) { return __builtin_log2f(x); }
GMA float fexp2(float x) { return __builtin_exp2f(x); }
GMA float frcp(float x) { return 1.0f / x; }
GMA float fsqrt(float x) { return __builtin_sqrtf(x); }
#endif

// a^(-2/3) for a > 0.
GMA float powm23(float a) { return fexp2(-0.6666666666666666f * flog2(a)); }

// FP64: a seed from the FP32 hardware log/exp where a is a normal float
// (relative error e below 1e-5: the log's absolute error grows with
// |log2 a| <= 126), then one Halley step on y^3 a^2 = 1: with t = a^2 y^3 =
// 1 + d, y (1 - d/3 + 2 d^2/9) leaves a relative error of about 5 e^3 (below
// 1e-15 for e < 5e-6; 4e-15 at the far ends of the float range, 1e-16 for
// the corner determinants of meshes in the unit range).  Other values (zero,
// subnormal, beyond the float range, inf, NaN) take the exact slow path.
GMA double powm23(double a)
{
    if(!(a >= 1.1754943508222875e-38 && a <= 3.4028234663852886e+38))
        return 1.0 / gm_pow23(a);
    const double y = (double)fexp2(-0.6666666666666666f * flog2((float)a));
    double t = a * y;
    t = t * t * y;                   // a^2 y^3
    const double d = t - 1.0;
    return __builtin_fma(y, d * __builtin_fma(d, 2.0 / 9.0, -1.0 / 3.0), y);
}

// |a|^(2/3) (the rare non-positive path).
GMA float pow23abs(float a) { return fexp2(0.6666666666666666f * flog2(__builtin_fabsf(a))); }
GMA double pow23abs(double a) { return gm_pow23(__builtin_fabs(a)); }

GMA float rcpT(float x) { return frcp(x); }
GMA double rcpT(double x) { return 1.0 / x; }

template<typename T>
GMA Corner<T> makeCorner(v3<T> c0, v3<T> c1, v3<T> c2, T r0, T r1, T r2)
{
    const v3<T> x12 = gmq::cross(c1, c2), x20 = gmq::cross(c2, c0), x01 = gmq::cross(c0, c1);
    Corner<T> k;
    k.d0 = gmq::dot(c0, x12);
    k.gx = r0 * x12.x + r1 * x20.x + r2 * x01.x;
    k.gy = r0 * x12.y + r1 * x20.y + r2 * x01.y;
    k.gz = r0 * x12.z + r1 * x20.z + r2 * x01.z;
    k.c0 = gmq::dot(c0, c0) + gmq::dot(c1, c1) + gmq::dot(c2, c2);
    k.hx = T(2) * (r0 * c0.x + r1 * c1.x + r2 * c2.x);
    k.hy = T(2) * (r0 * c0.y + r1 * c1.y + r2 * c2.y);
    k.hz = T(2) * (r0 * c0.z + r1 * c1.z + r2 * c2.z);
    k.s = r0 * r0 + r1 * r1 + r2 * r2;
    // The band edge is a heuristic: the hardware square root (within an ulp)
    // is precise enough.
    const float c0f = (float)k.c0;
    k.thr = NEAR_DEGENERATE * c0f * fsqrt(c0f);
    k.ref = 0;
    k.corner = 0;
    return k;
}

GMA float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
GMA double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Corner at offset x (|x|^2 = xx): d and f2 as FMA chains from the constant.
template<typename T> GMA T cornerDet(const Corner<T>& k, v3<T> x)
{
    return fmaT(k.gz, x.z, fmaT(k.gy, x.y, fmaT(k.gx, x.x, k.d0)));
}

template<typename T>
GMA void cornerAt(const Corner<T>& k, v3<T> x, T xx, T& d, T& f2)
{
    d = cornerDet<T>(k, x);
    f2 = fmaT(k.s, xx, fmaT(k.hz, x.z, fmaT(k.hy, x.y, fmaT(k.hx, x.x, k.c0))));
}

// 3 / q of a corner with d > 0 (the factor 3 is applied once per patch).
template<typename T> GMA T invQ3(T d, T f2) { return f2 * powm23(d); }

// q of a corner with d <= 0 (0 when d == 0 and f2 > 0, NaN when F == 0).
template<typename T> GMA T qNonPositive(T d, T f2)
{
    const T m = d < T(0) ? T(-3) : T(0);
    return m * pow23abs(d) * rcpT(f2);
}

// Patch accumulator: sum of 3/q over positive elements, smallest
// non-positive q, element count, and whether any element was non-positive.
template<typename T>
struct Acc
{
    T inv3 = T(0);
    T mn = T(0);
    int n = 0;
    bool bad = false;
    GMA void positive(T i3) { inv3 += i3; ++n; }
    GMA void nonPositive(T q)
    {
        // The first non-positive quality replaces the running value; later
        // ones keep the minimum (a NaN never replaces it).
        if(!bad) { mn = q; bad = true; }
        else if(q < mn) mn = q;
    }
    GMA T finish() const { return bad ? mn : (n ? T(3) * (T)n / inv3 : T(0)); }
};

// ---- Tetrahedra: one corner; columns e_i3 = v3 - v_i (metric-free
// segments), F = E Fr_TET_INV.  Moving vertex k: k == 3 adds x to every
// column of E, k < 3 subtracts it from column k; r = Fr^T a.
template<typename T>
GMA Corner<T> tetCorner(const v3<T> u[4], int k, uint32_t ref)
{
    const T R01 = T(-0.5773502691896257645091), R11 = T(1.154700538379251529018);
    const T R02 = T(-0.4082482904638630163662), R12 = T(-0.4082482904638630163662), R22 = T(1.224744871391589049099);
    const v3<T> e0 = u[3] - u[0], e1 = u[3] - u[1], e2 = u[3] - u[2];
    // F columns: f0 = e0, f1 = R01 e0 + R11 e1, f2 = R02 e0 + R12 e1 + R22 e2.
    const v3<T> f0 = e0;
    const v3<T> f1 = e0 * R01 + e1 * R11;
    const v3<T> f2 = e0 * R02 + e1 * R12 + e2 * R22;
    const T a0 = k == 3 ? T(1) : (k == 0 ? T(-1) : T(0));
    const T a1 = k == 3 ? T(1) : (k == 1 ? T(-1) : T(0));
    const T a2 = k == 3 ? T(1) : (k == 2 ? T(-1) : T(0));
    // r_j = sum_i a_i R_ij (R upper triangular, R00 = 1).
    const T r0 = a0;
    const T r1 = a0 * R01 + a1 * R11;
    const T r2 = a0 * R02 + a1 * R12 + a2 * R22;
    Corner<T> c = makeCorner<T>(f0, f1, f2, r0, r1, r2);
    c.ref = ref;
    return c;
}

// ---- Hexahedra: corner c has columns v[P] - v[M] for these (P, M) pairs
// (MeanRatioEvaluator::hexQuality corner frames), packed as 4 bits per
// vertex index: column j's P at bits 8j+4, M at 8j.
GMA uint32_t hexCornerCode(int c)
{
    switch(c)
    {
    case 0: return 0x034010u;  // (1,0) (4,0) (0,3)
    case 1: return 0x512110u;  // (1,0) (2,1) (5,1)
    case 2: return 0x236221u;  // (2,1) (6,2) (2,3)
    case 3: return 0x733230u;  // (3,0) (3,2) (7,3)
    case 4: return 0x745440u;  // (4,0) (5,4) (7,4)
    case 5: return 0x545651u;  // (5,1) (5,6) (5,4)
    case 6: return 0x766562u;  // (6,2) (6,5) (7,6)
    default: return 0x477673u; // (7,3) (7,6) (4,7)
    }
}

GMA void hexCornerPairs(int c, int j, int& P, int& M)
{
    const uint32_t code = hexCornerCode(c) >> (8 * j);
    P = (int)((code >> 4) & 15u);
    M = (int)(code & 15u);
}

// Staged hexahedron with moving vertex k: the 4 corners that involve k, and
// the 4 that do not folded into their 3/q sum (or their smallest quality
// when one is non-positive).  `fix(corner)` returns a constant corner's FP64
// determinant, used in FP32 when it is near-degenerate.
template<typename T>
struct alignas(16) Hex
{
    Corner<T> k[4];
    T constInv3;   // sum of 3/q over the constant corners (all positive)
    T constMin;    // smallest constant-corner q if one is non-positive, else +1
};

// Stages a hexahedron into h, which should live in memory (LDS): the slot of
// a moving corner depends on the vertex, and an indexed write to a register
// array would go through scratch.
template<typename T, class Fix> GMA void stageHex(Hex<T>& h, const v3<T> u[8], int k, uint32_t ref, const Fix& fix)
{
    int nv = 0;
    T inv3 = T(0), mn = T(1);
    bool bad = false;
#pragma unroll
    for(int c = 0; c < 8; ++c)
    {
        v3<T> col[3];
        T r[3];
        bool moves = false;
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
            int P, M;
            hexCornerPairs(c, j, P, M);
            col[j] = u[P] - u[M];
            r[j] = P == k ? T(1) : (M == k ? T(-1) : T(0));
            moves = moves || r[j] != T(0);
        }
        Corner<T> cc = makeCorner<T>(col[0], col[1], col[2], r[0], r[1], r[2]);
        cc.ref = ref;
        cc.corner = (uint32_t)c;
        if(moves)
        {
            if(nv < 4) h.k[nv] = cc;
            ++nv;
            continue;
        }
        if(sizeof(T) == 4 && __builtin_fabsf((float)cc.d0) < cc.thr) cc.d0 = (T)fix(cc);
        if(cc.d0 > T(0))
            inv3 += invQ3<T>(cc.d0, cc.c0);
        else
        {
            const T q = qNonPositive<T>(cc.d0, cc.c0);
            if(!bad || q < mn) mn = q;
            bad = true;
        }
    }
    // A non-positive constant corner makes the fast sum NaN (the exact rule
    // reads constMin only).
    h.constInv3 = bad ? T(__builtin_nan("")) : inv3;
    h.constMin = bad ? mn : T(1);
}

// Branch-free sums for the common case: 3/q of a corner, which is NaN
// (d < 0: the FP32 log of a negative number; FP64: explicit) or +inf (d == 0,
// or d so small that |d|^(-2/3) overflows) when the corner is not positive.
// A patch sum that is not finite (fastBad) sends the caller to the exact rule
// (addTet / addHex), which is rare; otherwise the patch quality is 3 n / sum.
template<typename T>
GMA T fastInv3(const Corner<T>& k, v3<T> x, T xx)
{
    T d, f2;
    cornerAt<T>(k, x, xx, d, f2);
    if constexpr(sizeof(T) == 4)
        return f2 * powm23(d);
    else
        return d > T(0) ? f2 * powm23(d) : T(__builtin_nan(""));
}

template<typename T> GMA T fastTet(const Corner<T>& k, v3<T> x, T xx) { return fastInv3<T>(k, x, xx); }

template<typename T>
GMA T fastHex(const Hex<T>& h, v3<T> x, T xx)
{
    T s = h.constInv3;
#pragma unroll
    for(int c = 0; c < 4; ++c) s += fastInv3<T>(h.k[c], x, xx);
    return T(0.125) * s;
}

template<typename T> GMA bool fastBad(T sum) { return !(sum < T(__builtin_huge_val())); }

// Adds one staged tetrahedron at offset x to the accumulator.
template<typename T>
GMA void addTet(Acc<T>& a, const Corner<T>& k, v3<T> x, T xx)
{
    T d, f2;
    cornerAt<T>(k, x, xx, d, f2);
    if(d > T(0)) a.positive(invQ3<T>(d, f2));
    else a.nonPositive(qNonPositive<T>(d, f2));
}

// Adds one staged hexahedron: its quality is 8 / sum(1/q_c) over its corners
// when all are positive (so 3/q_hex = (1/8) sum 3/q_c), else its smallest
// corner quality (MeanRatioEvaluator::hexQuality's mean-or-min).
template<typename T>
GMA void addHex(Acc<T>& a, const Hex<T>& h, v3<T> x, T xx)
{
    T inv3 = h.constInv3, mn = h.constMin;
    bool bad = h.constMin <= T(0);
#pragma unroll
    for(int c = 0; c < 4; ++c)
    {
        T d, f2;
        cornerAt<T>(h.k[c], x, xx, d, f2);
        if(d > T(0))
            inv3 += invQ3<T>(d, f2);
        else
        {
            const T q = qNonPositive<T>(d, f2);
            if(!bad || q < mn) mn = q;
            bad = true;
        }
    }
    if(bad) a.nonPositive(mn);
    else a.positive(T(0.125) * inv3);
}

#undef GMA

} // namespace gma

#endif // GM_AFFINE_H
