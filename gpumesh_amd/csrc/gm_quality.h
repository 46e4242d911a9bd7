// gm_quality.h -- element shape measures, measurers and metric samplers of the
// smoothing path, written once for both precisions (T = double or float).
//
// For T = double the operation order is the one of the reference's FP64 CPU
// code (each function cites it), so that a kernel built without contraction
// returns the same bits as a CPU build of the same order.  For T = float the
// formulas are identical but the transcendentals are the hardware ones.
//
// Matrices are column-major (c[col].row), as in the reference's vector library.
#ifndef GM_QUALITY_H
#define GM_QUALITY_H

#include "gm_libm.h"

#if !defined(__HIP_DEVICE_COMPILE__)
#include <cmath>
#endif

#if defined(__HIPCC__) || defined(__HIP__)
#define GMQ __host__ __device__ __forceinline__
#define GMQ_MEMBER __host__ __device__ __forceinline__
// A routine kept out of line: it is inlined at so many call sites (every
// Keast point of every corner) that the functions holding them grow long
// enough for the compiler to expand their branches through the register
// pair that holds a function's return address.
#define GMQ_OUTLINE static __host__ __device__ __attribute__((noinline))
#else
#define GMQ static inline
#define GMQ_MEMBER inline
#define GMQ_OUTLINE static __attribute__((noinline))
#endif

namespace gmq
{

template<typename T> struct v3
{
    T x, y, z;
};

template<typename T> GMQ v3<T> mk(T x, T y, T z) { v3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template<typename T> GMQ v3<T> operator+(v3<T> a, v3<T> b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template<typename T> GMQ v3<T> operator-(v3<T> a, v3<T> b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template<typename T> GMQ v3<T> operator-(v3<T> a) { return mk<T>(-a.x, -a.y, -a.z); }
template<typename T> GMQ v3<T> operator*(v3<T> a, T s) { return mk<T>(a.x * s, a.y * s, a.z * s); }
template<typename T> GMQ v3<T> operator*(T s, v3<T> a) { return mk<T>(s * a.x, s * a.y, s * a.z); }
template<typename T> GMQ v3<T> operator/(v3<T> a, T s) { return mk<T>(a.x / s, a.y / s, a.z / s); }
template<typename T> GMQ T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template<typename T> GMQ v3<T> cross(v3<T> x, v3<T> y)
{
    return mk<T>(x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y);
}
template<typename T> GMQ T tmin(T a, T b) { return (b < a) ? b : a; }
template<typename T> GMQ T tmax(T a, T b) { return (a < b) ? b : a; }
template<typename T> GMQ T tsign(T x) { return (T)(T(0) < x) - (T)(x < T(0)); }
GMQ double tabs(double x) { return __builtin_fabs(x); }
GMQ float tabs(float x) { return __builtin_fabsf(x); }

GMQ double tsqrt(double x) { return __builtin_sqrt(x); }
GMQ double tdiv(double a, double b) { return a / b; }
#if defined(__HIP_DEVICE_COMPILE__)
// FP32 on the device: the hardware square root and reciprocal (about 1 ulp;
// the FP32 modes are not bit-exact, FP64 keeps correctly rounded operations).
GMQ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
GMQ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#else
GMQ float tsqrt(float x) { return __builtin_sqrtf(x); }
GMQ float tdiv(float a, float b) { return a / b; }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
GMQ double tpow23(double x) { return gm_pow23(x); }
GMQ double tcbrt(double x) { return gm_cbrt_nonneg(x); }
GMQ double texp(double x) { return gm_exp(x); }
GMQ double tcos(double x) { return gm_cos(x); }
GMQ double tpow(double x, double y) { return gm_pow(x, y); }
#else
// Host code may take the C library's transcendentals (the reference's libm)
// instead of the project's, which the device always uses: the topologist
// does, so its greedy decisions are the ones the reference makes.  Per
// thread, off unless a caller turns it on.
inline thread_local bool hostLibc = false;
GMQ double tpow23(double x) { return hostLibc ? std::pow(x, 2.0 / 3.0) : gm_pow23(x); }
GMQ double tcbrt(double x) { return hostLibc ? std::pow(x, 1.0 / 3.0) : gm_cbrt_nonneg(x); }
GMQ double texp(double x) { return hostLibc ? std::exp(x) : gm_exp(x); }
GMQ double tcos(double x) { return hostLibc ? std::cos(x) : gm_cos(x); }
GMQ double tpow(double x, double y) { return hostLibc ? std::pow(x, y) : gm_pow(x, y); }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
GMQ float tpow23(float x) { return x > 0.0f ? __builtin_amdgcn_exp2f(0.6666666666666666f * __builtin_amdgcn_logf(x)) : (x == 0.0f ? 0.0f : __builtin_nanf("")); }
GMQ float tcbrt(float x) { return x > 0.0f ? __builtin_amdgcn_exp2f(0.3333333333333333f * __builtin_amdgcn_logf(x)) : (x == 0.0f ? 0.0f : __builtin_nanf("")); }
GMQ float texp(float x) { return __builtin_amdgcn_exp2f(1.4426950408889634f * x); }
#else
GMQ float tpow23(float x) { return (float)gm_pow23((double)x); }
GMQ float tcbrt(float x) { return (float)gm_cbrt_nonneg((double)x); }
GMQ float texp(float x) { return (float)gm_exp((double)x); }
#endif

template<typename T> struct m3
{
    v3<T> c[3];
};

template<typename T> GMQ m3<T> mcols(v3<T> a, v3<T> b, v3<T> d) { m3<T> r; r.c[0] = a; r.c[1] = b; r.c[2] = d; return r; }
template<typename T> GMQ m3<T> mdiag(T s)
{
    return mcols<T>(mk<T>(s, T(0), T(0)), mk<T>(T(0), s, T(0)), mk<T>(T(0), T(0), s));
}
template<typename T> GMQ T at(const m3<T>& m, int c, int r) { return r == 0 ? m.c[c].x : (r == 1 ? m.c[c].y : m.c[c].z); }

template<typename T> GMQ m3<T> mmul(const m3<T>& A, const m3<T>& B)
{
    m3<T> R;
#define GMQ_MM(j)                                                                  \
    R.c[j].x = A.c[0].x * at(B, j, 0) + A.c[1].x * at(B, j, 1) + A.c[2].x * at(B, j, 2); \
    R.c[j].y = A.c[0].y * at(B, j, 0) + A.c[1].y * at(B, j, 1) + A.c[2].y * at(B, j, 2); \
    R.c[j].z = A.c[0].z * at(B, j, 0) + A.c[1].z * at(B, j, 1) + A.c[2].z * at(B, j, 2);
    GMQ_MM(0) GMQ_MM(1) GMQ_MM(2)
#undef GMQ_MM
    return R;
}
template<typename T> GMQ v3<T> mvec(const m3<T>& m, v3<T> v)
{
    return mk<T>(m.c[0].x * v.x + m.c[1].x * v.y + m.c[2].x * v.z,
                 m.c[0].y * v.x + m.c[1].y * v.y + m.c[2].y * v.z,
                 m.c[0].z * v.x + m.c[1].z * v.y + m.c[2].z * v.z);
}
template<typename T> GMQ m3<T> madd(const m3<T>& A, const m3<T>& B) { return mcols<T>(A.c[0] + B.c[0], A.c[1] + B.c[1], A.c[2] + B.c[2]); }
template<typename T> GMQ m3<T> msub(const m3<T>& A, const m3<T>& B) { return mcols<T>(A.c[0] - B.c[0], A.c[1] - B.c[1], A.c[2] - B.c[2]); }
template<typename T> GMQ m3<T> mscale(const m3<T>& A, T s) { return mcols<T>(A.c[0] * s, A.c[1] * s, A.c[2] * s); }
template<typename T> GMQ m3<T> transpose(const m3<T>& m)
{
    return mcols<T>(mk<T>(m.c[0].x, m.c[1].x, m.c[2].x), mk<T>(m.c[0].y, m.c[1].y, m.c[2].y),
                    mk<T>(m.c[0].z, m.c[1].z, m.c[2].z));
}
template<typename T> GMQ T det(const m3<T>& m)
{
    return + m.c[0].x * (m.c[1].y * m.c[2].z - m.c[2].y * m.c[1].z)
           - m.c[1].x * (m.c[0].y * m.c[2].z - m.c[2].y * m.c[0].z)
           + m.c[2].x * (m.c[0].y * m.c[1].z - m.c[1].y * m.c[0].z);
}
template<typename T> GMQ m3<T> inverse(const m3<T>& m)
{
    T o = T(1) / det(m);
    m3<T> I;
    I.c[0].x = + (m.c[1].y * m.c[2].z - m.c[2].y * m.c[1].z) * o;
    I.c[1].x = - (m.c[1].x * m.c[2].z - m.c[2].x * m.c[1].z) * o;
    I.c[2].x = + (m.c[1].x * m.c[2].y - m.c[2].x * m.c[1].y) * o;
    I.c[0].y = - (m.c[0].y * m.c[2].z - m.c[2].y * m.c[0].z) * o;
    I.c[1].y = + (m.c[0].x * m.c[2].z - m.c[2].x * m.c[0].z) * o;
    I.c[2].y = - (m.c[0].x * m.c[2].y - m.c[2].x * m.c[0].y) * o;
    I.c[0].z = + (m.c[0].y * m.c[1].z - m.c[1].y * m.c[0].z) * o;
    I.c[1].z = - (m.c[0].x * m.c[1].z - m.c[1].x * m.c[0].z) * o;
    I.c[2].z = + (m.c[0].x * m.c[1].y - m.c[1].x * m.c[0].y) * o;
    return I;
}

// ---------------------------------------------------------------------------
// Crew (sampler / measurer / evaluator selection and parameters).
// ---------------------------------------------------------------------------
enum { EV_MEAN_RATIO = 0, EV_METRIC_CONFORMITY = 1 };
enum { SA_UNIFORM = 0, SA_SHELL = 1, SA_TEXTURE = 2, SA_ANALYTIC = 3 };

struct Crew
{
    double scaling, scaling2, scaling3;
    int tsize[3];
    double tmin[3], text[3];
    const float* texels;   // 6 per cell: m00 m01 m02 m11 m12 m22; cell (i,j,k) at ((k ny)+j) nx + i
    // The grid's cell size, its reciprocal, the box's low corner and the last
    // clamp corner in FP32, for the FP32 sampling (textureFields fills them).
    float tcellf[3], tinvf[3], tminf[3], tclampf[3];
    // The reference's analytic field: aspect ratio and the sampler's
    // rotation (column-major, rot[3 c + r]) and its transpose.
    double aspect;
    double rot[9], rotInv[9];
};

// UniformSampler::metricAt (Samplers/UniformSampler.cpp:45-51): K^2 I.
// Spherical-shell field (the anisotropic workload's analytic metric):
//   a = 1 + 9 exp(-50 (r - 0.3)^2),  M = K^2 (a^2 I + (1 - a^2) d d^T / r^2).
template<typename T> GMQ m3<T> shellMetric(const Crew& c, v3<T> p)
{
    T dx = p.x - T(0.5), dy = p.y - T(0.5), dz = p.z - T(0.5);
    T r2 = dx * dx + dy * dy + dz * dz;
    T r = tsqrt(r2);
    T u = r - T(0.3);
    T a = T(1) + T(9) * texp(T(-50) * (u * u));
    T a2 = a * a;
    T w = (r2 > T(0)) ? (T(1) - a2) / r2 : T(0);
    T k2 = (T)c.scaling2;
    T m00 = k2 * (a2 + w * (dx * dx));
    T m11 = k2 * (a2 + w * (dy * dy));
    T m22 = k2 * (a2 + w * (dz * dz));
    T m01 = k2 * (w * (dx * dy));
    T m02 = k2 * (w * (dx * dz));
    T m12 = k2 * (w * (dy * dz));
    return mcols<T>(mk<T>(m00, m01, m02), mk<T>(m01, m11, m12), mk<T>(m02, m12, m22));
}

template<typename T> GMQ m3<T> texel(const Crew& c, int i, int j, int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
    // The grid is in device memory: global-address-space loads, not flat ones.
    const __attribute__((address_space(1))) float* t =
        (const __attribute__((address_space(1))) float*)(const void*)c.texels +
        6 * (((long long)k * c.tsize[1] + j) * c.tsize[0] + i);
#else
    const float* t = c.texels + 6 * (((long long)k * c.tsize[1] + j) * c.tsize[0] + i);
#endif
    T m00 = (T)t[0], m01 = (T)t[1], m02 = (T)t[2], m11 = (T)t[3], m12 = (T)t[4], m22 = (T)t[5];
    return mcols<T>(mk<T>(m00, m01, m02), mk<T>(m01, m11, m12), mk<T>(m02, m12, m22));
}

template<typename T> GMQ m3<T> mixm(const m3<T>& x, const m3<T>& y, T a)
{
    return madd<T>(x, mcols<T>(a * (y.c[0] - x.c[0]), a * (y.c[1] - x.c[1]), a * (y.c[2] - x.c[2])));
}

GMQ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// TextureSampler::metricAt (Samplers/TextureSampler.cpp:450-488), cellId (:513-524).
// Fills the FP32 fields of a texture crew from tsize, tmin and text.
inline void textureFields(Crew& c)
{
    for(int k = 0; k < 3; ++k)
    {
        const double cell = c.text[k] / (double)c.tsize[k];
        c.tcellf[k] = (float)cell;
        c.tinvf[k] = (float)((double)c.tsize[k] / c.text[k]);
        c.tminf[k] = (float)c.tmin[k];
        c.tclampf[k] = (float)((c.tmin[k] + c.text[k]) - cell * 1.5);
    }
}

// TextureSampler::metricAt in FP32 (the FP32 passes): the same cell, clamp
// and weights, with the crew's FP32 cell size and reciprocal in place of the
// divisions.
GMQ m3<float> textureMetricF32(const Crew& c, v3<float> p)
{
    const float cx = c.tcellf[0], cy = c.tcellf[1], cz = c.tcellf[2];
    const float ix = c.tinvf[0], iy = c.tinvf[1], iz = c.tinvf[2];
    const v3<float> lo = mk<float>(c.tminf[0], c.tminf[1], c.tminf[2]);
    const v3<float> q = mk<float>(p.x - 0.5f * cx, p.y - 0.5f * cy, p.z - 0.5f * cz);
    const v3<float> cp0 = mk<float>(tmin(tmax(q.x, lo.x), c.tclampf[0]), tmin(tmax(q.y, lo.y), c.tclampf[1]),
                                    tmin(tmax(q.z, lo.z), c.tclampf[2]));
    const int i0 = clampi((int)((cp0.x - lo.x) * ix), 0, c.tsize[0] - 1);
    const int j0 = clampi((int)((cp0.y - lo.y) * iy), 0, c.tsize[1] - 1);
    const int k0 = clampi((int)((cp0.z - lo.z) * iz), 0, c.tsize[2] - 1);
    const int i1 = (i0 + 1 < c.tsize[0] - 1) ? i0 + 1 : c.tsize[0] - 1;
    const int j1 = (j0 + 1 < c.tsize[1] - 1) ? j0 + 1 : c.tsize[1] - 1;
    const int k1 = (k0 + 1 < c.tsize[2] - 1) ? k0 + 1 : c.tsize[2] - 1;
    const m3<float> m000 = texel<float>(c, i0, j0, k0), m100 = texel<float>(c, i1, j0, k0);
    const m3<float> m010 = texel<float>(c, i0, j1, k0), m110 = texel<float>(c, i1, j1, k0);
    const m3<float> m001 = texel<float>(c, i0, j0, k1), m101 = texel<float>(c, i1, j0, k1);
    const m3<float> m011 = texel<float>(c, i0, j1, k1), m111 = texel<float>(c, i1, j1, k1);
    v3<float> a = mk<float>((p.x - (lo.x + cx * ((float)i0 + 0.5f))) * ix, (p.y - (lo.y + cy * ((float)j0 + 0.5f))) * iy,
                            (p.z - (lo.z + cz * ((float)k0 + 0.5f))) * iz);
    a = mk<float>(tmin(tmax(a.x, 0.0f), 1.0f), tmin(tmax(a.y, 0.0f), 1.0f), tmin(tmax(a.z, 0.0f), 1.0f));
    const m3<float> mxy0 = mixm<float>(mixm<float>(m000, m100, a.x), mixm<float>(m010, m110, a.x), a.y);
    const m3<float> mxy1 = mixm<float>(mixm<float>(m001, m101, a.x), mixm<float>(m011, m111, a.x), a.y);
    return mixm<float>(mxy0, mxy1, a.z);
}

template<typename T> GMQ m3<T> textureMetric(const Crew& c, v3<T> position)
{
    if constexpr(sizeof(T) == 4)
        return textureMetricF32(c, position);
    v3<T> size = mk<T>((T)c.tsize[0], (T)c.tsize[1], (T)c.tsize[2]);
    v3<T> tminv = mk<T>((T)c.tmin[0], (T)c.tmin[1], (T)c.tmin[2]);
    v3<T> textv = mk<T>((T)c.text[0], (T)c.text[1], (T)c.text[2]);
    v3<T> cs = mk<T>(textv.x / size.x, textv.y / size.y, textv.z / size.z);
    v3<T> maxB = (tminv + textv) - (cs * T(1.5));
    v3<T> q = position + mk<T>(-cs.x, -cs.y, -cs.z) / T(2);
    v3<T> cp0 = mk<T>(tmin(tmax(q.x, tminv.x), maxB.x), tmin(tmax(q.y, tminv.y), maxB.y), tmin(tmax(q.z, tminv.z), maxB.z));
    v3<T> od = cp0 - tminv;
    v3<T> fid = mk<T>((od.x / textv.x) * size.x, (od.y / textv.y) * size.y, (od.z / textv.z) * size.z);
    int i0 = clampi((int)fid.x, 0, c.tsize[0] - 1), j0 = clampi((int)fid.y, 0, c.tsize[1] - 1), k0 = clampi((int)fid.z, 0, c.tsize[2] - 1);
    int i1 = (i0 + 1 < c.tsize[0] - 1) ? i0 + 1 : c.tsize[0] - 1;
    int j1 = (j0 + 1 < c.tsize[1] - 1) ? j0 + 1 : c.tsize[1] - 1;
    int k1 = (k0 + 1 < c.tsize[2] - 1) ? k0 + 1 : c.tsize[2] - 1;
    m3<T> m000 = texel<T>(c, i0, j0, k0), m100 = texel<T>(c, i1, j0, k0);
    m3<T> m010 = texel<T>(c, i0, j1, k0), m110 = texel<T>(c, i1, j1, k0);
    m3<T> m001 = texel<T>(c, i0, j0, k1), m101 = texel<T>(c, i1, j0, k1);
    m3<T> m011 = texel<T>(c, i0, j1, k1), m111 = texel<T>(c, i1, j1, k1);
    v3<T> c0c = mk<T>(cs.x * ((T)i0 + T(0.5)), cs.y * ((T)j0 + T(0.5)), cs.z * ((T)k0 + T(0.5)));
    v3<T> d = position - (tminv + c0c);
    v3<T> a = mk<T>(d.x / cs.x, d.y / cs.y, d.z / cs.z);
    a = mk<T>(tmin(tmax(a.x, T(0)), T(1)), tmin(tmax(a.y, T(0)), T(1)), tmin(tmax(a.z, T(0)), T(1)));
    m3<T> mx00 = mixm<T>(m000, m100, a.x);
    m3<T> mx10 = mixm<T>(m010, m110, a.x);
    m3<T> mx01 = mixm<T>(m001, m101, a.x);
    m3<T> mx11 = mixm<T>(m011, m111, a.x);
    m3<T> mxy0 = mixm<T>(mx00, mx10, a.y);
    m3<T> mxy1 = mixm<T>(mx01, mx11, a.y);
    return mixm<T>(mxy0, mxy1, a.z);
}

// AbstractSampler::vertMetric (Samplers/AbstractSampler.cpp:190-208), the
// AnalyticSampler's field (AnalyticSampler.cpp:45-49): along the rotated x
// axis the metric length scale is K a^(c^a) with c = (1 - cos(2.5 pi x'))/2,
// K across it; M = R^T diag(Mx, K^2, K^2) R.  Evaluated in FP64 in both
// precisions (the project's cos and pow, as on the device, unless a host
// caller selects the C library's).
GMQ_OUTLINE m3<double> referenceMetric(const Crew& c, v3<double> p)
{
    m3<double> R, RI;
    for(int j = 0; j < 3; ++j)
    {
        R.c[j] = mk<double>(c.rot[3 * j], c.rot[3 * j + 1], c.rot[3 * j + 2]);
        RI.c[j] = mk<double>(c.rotInv[3 * j], c.rotInv[3 * j + 1], c.rotInv[3 * j + 2]);
    }
    const double x = mvec<double>(R, p).x * (2.5 * 3.14159265358979323846);
    const double a = c.aspect;
    const double cc = (1.0 - tcos(x)) / 2.0;
    const double sizeX = c.scaling * tpow(a, tpow(cc, a));
    const double Mx = sizeX * sizeX;
    const double My = c.scaling2;
    const m3<double> M = mcols<double>(mk<double>(Mx, 0.0, 0.0), mk<double>(0.0, My, 0.0), mk<double>(0.0, 0.0, My));
    return mmul<double>(mmul<double>(RI, M), R);
}

template<typename T, int SA> GMQ m3<T> metricAt(const Crew& c, v3<T> p)
{
    if constexpr(SA == SA_ANALYTIC)
    {
        const m3<double> M = referenceMetric(c, mk<double>((double)p.x, (double)p.y, (double)p.z));
        m3<T> r;
        for(int j = 0; j < 3; ++j) r.c[j] = mk<T>((T)M.c[j].x, (T)M.c[j].y, (T)M.c[j].z);
        return r;
    }
    if(SA == SA_SHELL) return shellMetric<T>(c, p);
    if(SA == SA_TEXTURE) return textureMetric<T>(c, p);
    T s2 = (T)c.scaling * (T)c.scaling;
    return mdiag<T>(s2);
}

// A host build of these headers may define GMQ_WALK_STEP() to count the
// adaptive walk's iterations; the library leaves it empty.
#ifndef GMQ_WALK_STEP
#define GMQ_WALK_STEP()
#endif

// MetricWiseMeasurer::riemannianDistance, localized segment division
// (Measurers/MetricWiseMeasurer.cpp:65-116, DIFF_THRESHOLD :9).
template<typename T, int SA> GMQ T riemannianDistanceWise(const Crew& c, v3<T> a, v3<T> b)
{
    int curr = 0, base = 2;
    T len = T(0);
    v3<T> d = b - a;
    v3<T> bv = d / (T)base;
    while(curr < base)
    {
        GMQ_WALK_STEP();
        T p0 = ((T)curr + T(0.5)) / (T)base;
        T p1 = ((T)curr + T(1.5)) / (T)base;
        m3<T> M0 = metricAt<T, SA>(c, a + p0 * d);
        m3<T> M1 = metricAt<T, SA>(c, a + p1 * d);
        T l0 = tsqrt(dot(bv, mvec(M0, bv)));
        T l1 = tsqrt(dot(bv, mvec(M1, bv)));
        T sum = (l0 + l1);
        T diff = tabs(l0 - l1) / (sum / T(2));
        // A NaN difference (a zero-length pair, or a non-finite position or
        // metric) never falls below the threshold, so the subdivision would
        // run until the int indices overflow; the walk ends here instead.
        if(!(diff == diff))
            return len + sum;
        if(diff < T(0.01))
        {
            len += sum;
            curr += 2;
            if((curr & 2) == 0) { base >>= 1; curr >>= 1; }
            bv = d / (T)base;
        }
        else
        {
            base <<= 1;
            curr <<= 1;
            bv = bv / T(2);
        }
    }
    return len;
}

// MetricFreeMeasurer::riemannianSegment (MetricFreeMeasurer.cpp:36-43) for the
// uniform sampler; MetricWiseMeasurer::riemannianSegment (:118-126) otherwise.
template<typename T, int SA> GMQ v3<T> segment(const Crew& c, v3<T> a, v3<T> b)
{
    if(SA == SA_UNIFORM)
        return (b - a) * (T)c.scaling;
    v3<T> d = b - a;
    v3<T> n = d * (T(1) / tsqrt(dot(d, d)));
    return n * riemannianDistanceWise<T, SA>(c, a, b);
}

// Volumes (MetricFreeMeasurer.cpp:45-56, 77-105; MetricWiseMeasurer.cpp:128-139, 160-188).
template<typename T, int SA> GMQ T tetVolume(const Crew& c, const v3<T> vp[4])
{
    if(SA == SA_UNIFORM)
    {
        T d = det(mcols<T>(vp[3] - vp[0], vp[3] - vp[1], vp[3] - vp[2]));
        return (T)c.scaling3 * d / T(6);
    }
    T d = det(mcols<T>(segment<T, SA>(c, vp[0], vp[3]), segment<T, SA>(c, vp[1], vp[3]), segment<T, SA>(c, vp[2], vp[3])));
    return d / T(6);
}

// MetricFreeMeasurer::priVolume (:58-75) / MetricWiseMeasurer::priVolume (:141-158).
template<typename T, int SA> GMQ T priVolume(const Crew& c, const v3<T> vp[6])
{
    v3<T> e02, e12, e32, e42, e52;
    if(SA == SA_UNIFORM)
    {
        e02 = vp[2] - vp[0]; e12 = vp[2] - vp[1]; e32 = vp[2] - vp[3]; e42 = vp[2] - vp[4]; e52 = vp[2] - vp[5];
    }
    else
    {
        e02 = segment<T, SA>(c, vp[0], vp[2]); e12 = segment<T, SA>(c, vp[1], vp[2]);
        e32 = segment<T, SA>(c, vp[3], vp[2]); e42 = segment<T, SA>(c, vp[4], vp[2]);
        e52 = segment<T, SA>(c, vp[5], vp[2]);
    }
    T s = T(0);
    s += det(mcols<T>(e32, e52, e42));
    s += det(mcols<T>(e02, e32, e42));
    s += det(mcols<T>(e12, e02, e42));
    if(SA == SA_UNIFORM)
        return (T)c.scaling3 * s / T(6);
    return s / T(6);
}

template<typename T, int SA> GMQ T hexVolume(const Crew& c, const v3<T> vp[8])
{
    T s = T(0);
    if(SA == SA_UNIFORM)
    {
        s += det(mcols<T>(vp[0] - vp[1], vp[0] - vp[4], vp[0] - vp[3]));
        s += det(mcols<T>(vp[2] - vp[1], vp[2] - vp[3], vp[2] - vp[6]));
        s += det(mcols<T>(vp[5] - vp[4], vp[5] - vp[1], vp[5] - vp[6]));
        s += det(mcols<T>(vp[7] - vp[4], vp[7] - vp[6], vp[7] - vp[3]));
        s += det(mcols<T>(vp[4] - vp[1], vp[4] - vp[6], vp[4] - vp[3]));
        return (T)c.scaling3 * s / T(6);
    }
    s += det(mcols<T>(segment<T, SA>(c, vp[1], vp[0]), segment<T, SA>(c, vp[4], vp[0]), segment<T, SA>(c, vp[3], vp[0])));
    s += det(mcols<T>(segment<T, SA>(c, vp[1], vp[2]), segment<T, SA>(c, vp[3], vp[2]), segment<T, SA>(c, vp[6], vp[2])));
    s += det(mcols<T>(segment<T, SA>(c, vp[4], vp[5]), segment<T, SA>(c, vp[1], vp[5]), segment<T, SA>(c, vp[6], vp[5])));
    s += det(mcols<T>(segment<T, SA>(c, vp[4], vp[7]), segment<T, SA>(c, vp[6], vp[7]), segment<T, SA>(c, vp[3], vp[7])));
    s += det(mcols<T>(segment<T, SA>(c, vp[1], vp[4]), segment<T, SA>(c, vp[6], vp[4]), segment<T, SA>(c, vp[3], vp[4])));
    return s / T(6);
}

// Fr_TET_INV (Evaluators/AbstractEvaluator.cpp:28-31).
template<typename T> GMQ m3<T> frTetInv()
{
    return mcols<T>(mk<T>(T(1), T(0), T(0)),
                    mk<T>(T(-0.5773502691896257645091), T(1.154700538379251529018), T(0)),
                    mk<T>(T(-0.4082482904638630163662), T(-0.4082482904638630163662), T(1.224744871391589049099)));
}

// Fr_PRI_INV (Evaluators/AbstractEvaluator.cpp:32-35).
template<typename T> GMQ m3<T> frPriInv()
{
    return mcols<T>(mk<T>(T(1), T(0), T(0)),
                    mk<T>(T(-0.5773502691896257645091), T(1.154700538379251529018), T(0)),
                    mk<T>(T(0), T(0), T(1)));
}

// MeanRatioEvaluator::cornerQuality (MeanRatioEvaluator.cpp:22-32).
template<typename T> GMQ T cornerQuality(const m3<T>& F)
{
    T d = det(F);
    T f2 = dot(F.c[0], F.c[0]) + dot(F.c[1], F.c[1]) + dot(F.c[2], F.c[2]);
    if constexpr(sizeof(T) == 4)
    {
        // FP32: a near-degenerate corner (|d| < 1e-4 f2^(3/2), quality below
        // about 0.006) gets its determinant, and so its sign, in FP64.
        if(d * d < T(1e-8) * f2 * f2 * f2)
        {
            m3<double> Fd;
            for(int j = 0; j < 3; ++j) Fd.c[j] = mk<double>((double)F.c[j].x, (double)F.c[j].y, (double)F.c[j].z);
            d = (T)det(Fd);
        }
    }
    return tdiv(tsign(d) * T(3) * tpow23(tabs(d)), f2);
}

// MetricConformityEvaluator::specifiedMetric (:25-45): 5-point Keast rule.
template<typename T, int SA> GMQ m3<T> specifiedMetric(const Crew& c, v3<T> v0, v3<T> v1, v3<T> v2, v3<T> v3_)
{
    const T H = T(0.5);
    const T Q = T((1.0 - 0.5) / 3.0);
    m3<T> r = mscale<T>(metricAt<T, SA>(c, (v0 + v1 + v2 + v3_) / T(4)), T(-0.8));
    r = madd<T>(r, mscale<T>(metricAt<T, SA>(c, v0 * H + v1 * Q + v2 * Q + v3_ * Q), T(0.45)));
    r = madd<T>(r, mscale<T>(metricAt<T, SA>(c, v0 * Q + v1 * H + v2 * Q + v3_ * Q), T(0.45)));
    r = madd<T>(r, mscale<T>(metricAt<T, SA>(c, v0 * Q + v1 * Q + v2 * H + v3_ * Q), T(0.45)));
    r = madd<T>(r, mscale<T>(metricAt<T, SA>(c, v0 * Q + v1 * Q + v2 * Q + v3_ * H), T(0.45)));
    return r;
}

// MetricConformityEvaluator::metricConformity (:47-72).
template<typename T> GMQ T metricConformity(const m3<T>& Fk, const m3<T>& Ms)
{
    m3<T> Fi = inverse<T>(transpose<T>(Fk));
    m3<T> Mk = mmul<T>(Fi, transpose<T>(Fi));
    m3<T> Mki = inverse<T>(Mk);
    m3<T> Msi = inverse<T>(Ms);
    m3<T> t = msub<T>(madd<T>(mmul<T>(Mki, Ms), mmul<T>(Msi, Mk)), mdiag<T>(T(2)));
    T f2 = dot(t.c[0], t.c[0]) + dot(t.c[1], t.c[1]) + dot(t.c[2], t.c[2]);
    if(f2 == f2)
        return tsign(det(Fk)) / tsqrt(T(1) + tsqrt(f2));
    return T(0);
}

// Tet quality: MeanRatioEvaluator::tetQuality (:34-50) or
// MetricConformityEvaluator::tetQuality (:74-91).
template<typename T, int EV, int SA> GMQ T tetQuality(const Crew& c, const v3<T> vp[4])
{
    if(EV == EV_METRIC_CONFORMITY)
    {
        m3<T> Fk = mmul<T>(mcols<T>(vp[3] - vp[0], vp[3] - vp[1], vp[3] - vp[2]), frTetInv<T>());
        return metricConformity<T>(Fk, specifiedMetric<T, SA>(c, vp[0], vp[1], vp[2], vp[3]));
    }
    v3<T> e03 = segment<T, SA>(c, vp[0], vp[3]);
    v3<T> e13 = segment<T, SA>(c, vp[1], vp[3]);
    v3<T> e23 = segment<T, SA>(c, vp[2], vp[3]);
    return cornerQuality<T>(mmul<T>(mcols<T>(e03, e13, e23), frTetInv<T>()));
}

template<typename T> GMQ T hexMeanOrMin(const T q[8])
{
    if(q[0] <= T(0) || q[1] <= T(0) || q[2] <= T(0) || q[3] <= T(0) ||
       q[4] <= T(0) || q[5] <= T(0) || q[6] <= T(0) || q[7] <= T(0))
    {
        T mn = q[0];
        for(int k = 1; k < 8; ++k) if(q[k] < mn) mn = q[k];
        return mn;
    }
    return T(8) / (T(1) / q[0] + T(1) / q[1] + T(1) / q[2] + T(1) / q[3] +
                   T(1) / q[4] + T(1) / q[5] + T(1) / q[6] + T(1) / q[7]);
}

// The prism's mean-or-min rule (MeanRatioEvaluator.cpp:85-103,
// MetricConformityEvaluator.cpp:133-151).
template<typename T> GMQ T priMeanOrMin(const T q[6])
{
    if(q[0] <= T(0) || q[1] <= T(0) || q[2] <= T(0) || q[3] <= T(0) || q[4] <= T(0) || q[5] <= T(0))
    {
        T mn = q[0];
        for(int k = 1; k < 6; ++k) if(q[k] < mn) mn = q[k];
        return mn;
    }
    return T(6) / (T(1) / q[0] + T(1) / q[1] + T(1) / q[2] + T(1) / q[3] + T(1) / q[4] + T(1) / q[5]);
}

// Prism quality: MeanRatioEvaluator::priQuality (:52-104) or
// MetricConformityEvaluator::priQuality (:93-152).  The corner frames
// (-e01, e20, e03) ... times Fr_PRI_INV are not invariant under edge swaps.
template<typename T, int EV, int SA> GMQ T priQuality(const Crew& c, const v3<T> vp[6])
{
    v3<T> e03, e14, e25, e01, e12, e20, e34, e45, e53;
    if(EV == EV_METRIC_CONFORMITY)
    {
        e03 = vp[3] - vp[0]; e14 = vp[4] - vp[1]; e25 = vp[5] - vp[2];
        e01 = vp[1] - vp[0]; e12 = vp[2] - vp[1]; e20 = vp[0] - vp[2];
        e34 = vp[4] - vp[3]; e45 = vp[5] - vp[4]; e53 = vp[3] - vp[5];
    }
    else
    {
        e03 = segment<T, SA>(c, vp[0], vp[3]); e14 = segment<T, SA>(c, vp[1], vp[4]);
        e25 = segment<T, SA>(c, vp[2], vp[5]); e01 = segment<T, SA>(c, vp[0], vp[1]);
        e12 = segment<T, SA>(c, vp[1], vp[2]); e20 = segment<T, SA>(c, vp[2], vp[0]);
        e34 = segment<T, SA>(c, vp[3], vp[4]); e45 = segment<T, SA>(c, vp[4], vp[5]);
        e53 = segment<T, SA>(c, vp[5], vp[3]);
    }
    const m3<T> fr = frPriInv<T>();
    const m3<T> F0 = mmul<T>(mcols<T>(-e01, e20, e03), fr), F1 = mmul<T>(mcols<T>(-e12, e01, e14), fr);
    const m3<T> F2 = mmul<T>(mcols<T>(-e20, e12, e25), fr), F3 = mmul<T>(mcols<T>(-e34, e53, e03), fr);
    const m3<T> F4 = mmul<T>(mcols<T>(-e45, e34, e14), fr), F5 = mmul<T>(mcols<T>(-e53, e45, e25), fr);
    T q[6];
    if(EV == EV_METRIC_CONFORMITY)
    {
        q[0] = metricConformity<T>(F0, specifiedMetric<T, SA>(c, vp[0], vp[1], vp[2], vp[3]));
        q[1] = metricConformity<T>(F1, specifiedMetric<T, SA>(c, vp[0], vp[1], vp[2], vp[4]));
        q[2] = metricConformity<T>(F2, specifiedMetric<T, SA>(c, vp[0], vp[1], vp[2], vp[5]));
        q[3] = metricConformity<T>(F3, specifiedMetric<T, SA>(c, vp[0], vp[3], vp[4], vp[5]));
        q[4] = metricConformity<T>(F4, specifiedMetric<T, SA>(c, vp[1], vp[3], vp[4], vp[5]));
        q[5] = metricConformity<T>(F5, specifiedMetric<T, SA>(c, vp[2], vp[3], vp[4], vp[5]));
    }
    else
    {
        q[0] = cornerQuality<T>(F0); q[1] = cornerQuality<T>(F1); q[2] = cornerQuality<T>(F2);
        q[3] = cornerQuality<T>(F3); q[4] = cornerQuality<T>(F4); q[5] = cornerQuality<T>(F5);
    }
    return priMeanOrMin<T>(q);
}

// Hex quality: MeanRatioEvaluator::hexQuality (:106-167) or
// MetricConformityEvaluator::hexQuality (:154-224).
template<typename T, int EV, int SA> GMQ T hexQuality(const Crew& c, const v3<T> vp[8])
{
    T q[8];
    v3<T> e01, e03, e04, e12, e15, e23, e26, e37, e45, e47, e56, e67;
    if(EV == EV_METRIC_CONFORMITY)
    {
        e01 = vp[1] - vp[0]; e03 = vp[3] - vp[0]; e04 = vp[4] - vp[0]; e12 = vp[2] - vp[1];
        e15 = vp[5] - vp[1]; e23 = vp[3] - vp[2]; e26 = vp[6] - vp[2]; e37 = vp[7] - vp[3];
        e45 = vp[5] - vp[4]; e47 = vp[7] - vp[4]; e56 = vp[6] - vp[5]; e67 = vp[7] - vp[6];
        q[0] = metricConformity<T>(mcols<T>(e01, e04, -e03), specifiedMetric<T, SA>(c, vp[0], vp[1], vp[3], vp[4]));
        q[1] = metricConformity<T>(mcols<T>(e01, e12, e15), specifiedMetric<T, SA>(c, vp[0], vp[1], vp[2], vp[5]));
        q[2] = metricConformity<T>(mcols<T>(e12, e26, -e23), specifiedMetric<T, SA>(c, vp[1], vp[2], vp[3], vp[6]));
        q[3] = metricConformity<T>(mcols<T>(e03, e23, e37), specifiedMetric<T, SA>(c, vp[0], vp[2], vp[3], vp[7]));
        q[4] = metricConformity<T>(mcols<T>(e04, e45, e47), specifiedMetric<T, SA>(c, vp[0], vp[4], vp[5], vp[7]));
        q[5] = metricConformity<T>(mcols<T>(e15, -e56, e45), specifiedMetric<T, SA>(c, vp[1], vp[4], vp[5], vp[6]));
        q[6] = metricConformity<T>(mcols<T>(e26, e56, e67), specifiedMetric<T, SA>(c, vp[2], vp[5], vp[6], vp[7]));
        q[7] = metricConformity<T>(mcols<T>(e37, e67, -e47), specifiedMetric<T, SA>(c, vp[3], vp[4], vp[6], vp[7]));
        return hexMeanOrMin<T>(q);
    }
    e01 = segment<T, SA>(c, vp[0], vp[1]); e03 = segment<T, SA>(c, vp[0], vp[3]);
    e04 = segment<T, SA>(c, vp[0], vp[4]); e12 = segment<T, SA>(c, vp[1], vp[2]);
    e15 = segment<T, SA>(c, vp[1], vp[5]); e23 = segment<T, SA>(c, vp[2], vp[3]);
    e26 = segment<T, SA>(c, vp[2], vp[6]); e37 = segment<T, SA>(c, vp[3], vp[7]);
    e45 = segment<T, SA>(c, vp[4], vp[5]); e47 = segment<T, SA>(c, vp[4], vp[7]);
    e56 = segment<T, SA>(c, vp[5], vp[6]); e67 = segment<T, SA>(c, vp[6], vp[7]);
    q[0] = cornerQuality<T>(mcols<T>(e01, e04, -e03));
    q[1] = cornerQuality<T>(mcols<T>(e01, e12, e15));
    q[2] = cornerQuality<T>(mcols<T>(e12, e26, -e23));
    q[3] = cornerQuality<T>(mcols<T>(e03, e23, e37));
    q[4] = cornerQuality<T>(mcols<T>(e04, e45, e47));
    q[5] = cornerQuality<T>(mcols<T>(e15, -e56, e45));
    q[6] = cornerQuality<T>(mcols<T>(e26, e56, e67));
    q[7] = cornerQuality<T>(mcols<T>(e37, e67, -e47));
    return hexMeanOrMin<T>(q);
}

// AbstractEvaluator::accumulatePatchQuality / finalizePatchQuality
// (Evaluators/AbstractEvaluator.cpp:778-803).
template<typename T> GMQ void accumulatePatch(T& pq, T& pw, T q)
{
    if(pw != T(-1) && q > T(0)) { pw += T(1); pq += T(1) / q; }
    else { pw = T(-1); pq = tmin(pq, q); }
}
template<typename T> GMQ T finalizePatch(T pq, T pw) { return (pw > T(0)) ? pw / pq : pq; }

} // namespace gmq

#endif // GM_QUALITY_H
