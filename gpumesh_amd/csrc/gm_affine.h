// gm_affine.h -- FP32 patch evaluation of the mean-ratio evaluator with the
// metric-free measurer, for optimisers that try many positions of one vertex
// while the rest of its patch stays put.
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
// position, turns a corner evaluation into a handful of FMAs, one log and one
// exp.  The identity is exact; the results differ from the direct formula
// only by FP32 rounding.  Mean ratio is invariant under uniform scaling, so
// the metric scaling factor drops out.
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

// One corner: d0, g (3), c0, h (3), s, then the near-degenerate threshold
// and the corner's identity (element reference bits, hex corner index) for
// the FP64 determinant recheck; three 16-byte rows, read with three vector
// LDS loads.
struct alignas(16) Corner
{
    float d0, gx, gy, gz, c0, hx, hy, hz, s, thr, ref, corner;
};

// A corner whose FP32 determinant is below thr = NEAR_DEGENERATE * c0^(3/2)
// in magnitude (quality below about 0.006) is near-degenerate: its sign --
// valid or inverted -- is decided in FP64 from the FP32 positions (for the
// constant corners of a staged hexahedron, and for every corner at a
// vertex's chosen new position).  FP32 rounding of d is about 1e-7 c0^(3/2),
// far inside the band.
constexpr float NEAR_DEGENERATE = 1.0e-4f;

#define GMA __host__ __device__ inline

GMA Corner makeCorner(v3<float> c0, v3<float> c1, v3<float> c2, float r0, float r1, float r2)
{
    const v3<float> x12 = gmq::cross(c1, c2), x20 = gmq::cross(c2, c0), x01 = gmq::cross(c0, c1);
    Corner k;
    k.d0 = gmq::dot(c0, x12);
    k.gx = r0 * x12.x + r1 * x20.x + r2 * x01.x;
    k.gy = r0 * x12.y + r1 * x20.y + r2 * x01.y;
    k.gz = r0 * x12.z + r1 * x20.z + r2 * x01.z;
    k.c0 = gmq::dot(c0, c0) + gmq::dot(c1, c1) + gmq::dot(c2, c2);
    k.hx = 2.0f * (r0 * c0.x + r1 * c1.x + r2 * c2.x);
    k.hy = 2.0f * (r0 * c0.y + r1 * c1.y + r2 * c2.y);
    k.hz = 2.0f * (r0 * c0.z + r1 * c1.z + r2 * c2.z);
    k.s = r0 * r0 + r1 * r1 + r2 * r2;
    k.thr = NEAR_DEGENERATE * k.c0 * __builtin_sqrtf(k.c0);
    k.ref = 0.0f;
    k.corner = 0.0f;
    return k;
}

#if defined(__HIP_DEVICE_COMPILE__)
GMA float flog2(float x) { return __builtin_amdgcn_logf(x); }
GMA float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
GMA float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
#else
GMA float flog2(float x) { return __builtin_log2f(x); }
GMA float fexp2(float x) { return __builtin_exp2f(x); }
GMA float frcp(float x) { return 1.0f / x; }
#endif
GMA float asFloat(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
GMA uint32_t asBits(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }

// Corner at offset x (|x|^2 = xx): returns d and f2.
GMA void cornerAt(const Corner& k, v3<float> x, float xx, float& d, float& f2)
{
    d = k.d0 + (k.gx * x.x + k.gy * x.y + k.gz * x.z);
    f2 = k.c0 + (k.hx * x.x + k.hy * x.y + k.hz * x.z) + k.s * xx;
}

// 3 / q of a corner with d > 0 (the factor 3 is applied once per patch).
GMA float invQ3(float d, float f2) { return f2 * fexp2(-0.6666666666666666f * flog2(d)); }

// q of a corner with d <= 0 (0 when d == 0 and f2 > 0, NaN when F == 0).
GMA float qNonPositive(float d, float f2)
{
    const float m = d < 0.0f ? -3.0f : 0.0f;
    return m * fexp2(0.6666666666666666f * flog2(-d)) * frcp(f2);
}

// Patch accumulator: sum of 3/q over positive elements, smallest
// non-positive q, element count, and whether any element was non-positive.
struct Acc
{
    float inv3 = 0.0f;
    float mn = 0.0f;
    int n = 0;
    bool bad = false;
    GMA void positive(float i3) { inv3 += i3; ++n; }
    GMA void nonPositive(float q)
    {
        // The first non-positive quality replaces the running value; later
        // ones keep the minimum (a NaN never replaces it).
        if(!bad) { mn = q; bad = true; }
        else if(q < mn) mn = q;
    }
    GMA float finish() const { return bad ? mn : (n ? 3.0f * (float)n / inv3 : 0.0f); }
};

// ---- Tetrahedra: one corner; columns e_i3 = v3 - v_i (metric-free
// segments), F = E Fr_TET_INV.  Moving vertex k: k == 3 adds x to every
// column of E, k < 3 subtracts it from column k; r = Fr^T a.
GMA Corner tetCorner(const v3<float> u[4], int k, uint32_t ref)
{
    const float R01 = -0.5773502691896257645091f, R11 = 1.154700538379251529018f;
    const float R02 = -0.4082482904638630163662f, R12 = -0.4082482904638630163662f, R22 = 1.224744871391589049099f;
    const v3<float> e0 = u[3] - u[0], e1 = u[3] - u[1], e2 = u[3] - u[2];
    // F columns: f0 = e0, f1 = R01 e0 + R11 e1, f2 = R02 e0 + R12 e1 + R22 e2.
    const v3<float> f0 = e0;
    const v3<float> f1 = e0 * R01 + e1 * R11;
    const v3<float> f2 = e0 * R02 + e1 * R12 + e2 * R22;
    const float a0 = k == 3 ? 1.0f : (k == 0 ? -1.0f : 0.0f);
    const float a1 = k == 3 ? 1.0f : (k == 1 ? -1.0f : 0.0f);
    const float a2 = k == 3 ? 1.0f : (k == 2 ? -1.0f : 0.0f);
    // r_j = sum_i a_i R_ij (R upper triangular, R00 = 1).
    const float r0 = a0;
    const float r1 = a0 * R01 + a1 * R11;
    const float r2 = a0 * R02 + a1 * R12 + a2 * R22;
    Corner c = makeCorner(f0, f1, f2, r0, r1, r2);
    c.ref = asFloat(ref);
    return c;
}

// ---- Hexahedra: corner c has columns v[P] - v[M] for the (P, M) pairs
// below (MeanRatioEvaluator::hexQuality corner frames).
// Packed as 4 bits per vertex index: column j's P at bits 8j+4, M at 8j.
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
// determinant, used when it is near-degenerate.
struct alignas(16) Hex
{
    Corner k[4];
    float constInv3;   // sum of 3/q over the constant corners (all positive)
    float constMin;    // smallest constant-corner q if one is non-positive, else +1
    float pad[2];
};

template<class Fix> GMA Hex stageHex(const v3<float> u[8], int k, uint32_t ref, const Fix& fix)
{
    Hex h;
    int nv = 0;
    float inv3 = 0.0f, mn = 1.0f;
    bool bad = false;
#pragma unroll
    for(int c = 0; c < 8; ++c)
    {
        v3<float> col[3];
        float r[3];
        bool moves = false;
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
            int P, M;
            hexCornerPairs(c, j, P, M);
            col[j] = u[P] - u[M];
            r[j] = P == k ? 1.0f : (M == k ? -1.0f : 0.0f);
            moves = moves || r[j] != 0.0f;
        }
        Corner cc = makeCorner(col[0], col[1], col[2], r[0], r[1], r[2]);
        cc.ref = asFloat(ref);
        cc.corner = asFloat((uint32_t)c);
        if(moves)
        {
            if(nv < 4) h.k[nv] = cc;
            ++nv;
            continue;
        }
        if(__builtin_fabsf(cc.d0) < cc.thr) cc.d0 = fix(cc);
        if(cc.d0 > 0.0f)
            inv3 += invQ3(cc.d0, cc.c0);
        else
        {
            const float q = qNonPositive(cc.d0, cc.c0);
            if(!bad || q < mn) mn = q;
            bad = true;
        }
    }
    h.constInv3 = inv3;
    h.constMin = bad ? mn : 1.0f;
    h.pad[0] = h.pad[1] = 0.0f;
    return h;
}

// Branch-free sums for the common case: 3/q of a corner assuming d > 0, and
// whether that assumption failed.  A patch whose elements are all positive
// has quality 3 n / sum; otherwise the caller evaluates it again with the
// exact rule (addTet / addHex), which is rare.
GMA float fastInv3(const Corner& k, v3<float> x, float xx, bool& bad)
{
    float d, f2;
    cornerAt(k, x, xx, d, f2);
    bad = bad || !(d > 0.0f);
    return f2 * fexp2(-0.6666666666666666f * flog2(__builtin_fabsf(d)));
}

GMA float fastTet(const Corner& k, v3<float> x, float xx, bool& bad) { return fastInv3(k, x, xx, bad); }

GMA float fastHex(const Hex& h, v3<float> x, float xx, bool& bad)
{
    float s = h.constInv3;
    bad = bad || !(h.constMin > 0.0f);
#pragma unroll
    for(int c = 0; c < 4; ++c) s += fastInv3(h.k[c], x, xx, bad);
    return 0.125f * s;
}

// Adds one staged tetrahedron at offset x to the accumulator.
GMA void addTet(Acc& a, const Corner& k, v3<float> x, float xx)
{
    float d, f2;
    cornerAt(k, x, xx, d, f2);
    if(d > 0.0f) a.positive(invQ3(d, f2));
    else a.nonPositive(qNonPositive(d, f2));
}

// Adds one staged hexahedron: its quality is 8 / sum(1/q_c) over its corners
// when all are positive (so 3/q_hex = (1/8) sum 3/q_c), else its smallest
// corner quality (MeanRatioEvaluator::hexQuality's mean-or-min).
GMA void addHex(Acc& a, const Hex& h, v3<float> x, float xx)
{
    float inv3 = h.constInv3, mn = h.constMin;
    bool bad = h.constMin <= 0.0f;
#pragma unroll
    for(int c = 0; c < 4; ++c)
    {
        float d, f2;
        cornerAt(h.k[c], x, xx, d, f2);
        if(d > 0.0f)
            inv3 += invQ3(d, f2);
        else
        {
            const float q = qNonPositive(d, f2);
            if(!bad || q < mn) mn = q;
            bad = true;
        }
    }
    if(bad) a.nonPositive(mn);
    else a.positive(0.125f * inv3);
}

#undef GMA

} // namespace gma

#endif // GM_AFFINE_H
