// gm_kernels.h -- kernel bodies of the smoothing path, instantiated once per
// (precision, evaluator, sampler) by kernels_inst.hip: FP64 objects are built
// without contraction (bit-comparable to an FP64 CPU build of the same
// operation order), FP32 objects with it.
#ifndef GM_KERNELS_H
#define GM_KERNELS_H

#include "gm_device.h"
#include "gm_affine.h"
#include "gm_reduce.h"
#include "gm_smoothers.h"

#include <type_traits>

namespace gmd
{
using gmq::v3;
using gmq::mk;

template<typename T> struct PosOf;
template<> struct PosOf<double>
{
    typedef double4 type;
    static __device__ __forceinline__ double4* get(const DevMesh& m) { return m.pos64; }
    static double4* host(const DevMesh& m) { return m.pos64; }
};
template<> struct PosOf<float>
{
    typedef float4 type;
    static __device__ __forceinline__ float4* get(const DevMesh& m) { return m.pos32; }
    static float4* host(const DevMesh& m) { return m.pos32; }
};

// Work item i of a vertex launch: (mesh index, vertex slot), slot NO_ITEM for
// a hole.  A colour group occupies consecutive slots, and its span holds the
// first one.
__device__ __forceinline__ uint2 workItem(const DevMesh* meshes, const uint4* segs, uint32_t nsegs, uint32_t i)
{
    (void)meshes;
    const uint4 s = segs[spanOf(segs, nsegs, i)];
    const uint32_t r = i - s.z;
    return make_uint2(s.x, r < s.w ? s.y + r : NO_ITEM);
}

// The same for a kernel whose wavefront's items lie in one ITEM_ALIGN window
// (wave and lane kernels): the span is found once per wavefront from its
// first active lane, so the search, the span and the mesh descriptor are
// wavefront-uniform (scalar loads).
__device__ __forceinline__ uint4 waveSpan(const uint4* segs, uint32_t nsegs, uint32_t i)
{
    const uint32_t first = __builtin_amdgcn_readfirstlane(i);
    const uint32_t k = __builtin_amdgcn_readfirstlane(spanOf(segs, nsegs, first));
    return segs[k];
}

__device__ __forceinline__ uint32_t spanVertex(const DevMesh& m, const uint4& s, uint32_t i)
{
    (void)m;
    const uint32_t r = i - s.z;
    return r < s.w ? s.y + r : NO_ITEM;
}

// The hardware hands workgroup b of a launch to XCD b mod 8, and each XCD
// has its own L2.  Vertex launches take their work in this order instead:
// XCD x runs a contiguous run of about 1/8 of the launch's work items, so
// neighbouring vertices, whose patches share vertices and cache lines, meet
// in one L2.  A bijection of [0, gridDim.x).
__device__ __forceinline__ uint32_t xcdBlock()
{
    const uint32_t b = blockIdx.x, n = gridDim.x;
    const uint32_t q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Block b of a block launch: (mesh index, block index within the mesh).
__device__ __forceinline__ uint2 blockItem(const uint4* spans, uint32_t nspans, uint32_t b)
{
    const uint4 s = spans[spanOf(spans, nspans, b)];
    return make_uint2(s.x, b - s.z);
}

template<typename T, typename P> __device__ __forceinline__ v3<T> ld(const P* pos, uint32_t v)
{
    P q = pos[v];
    return mk<T>((T)q.x, (T)q.y, (T)q.z);
}

template<typename P, typename T> __device__ __forceinline__ void st(P* pos, uint32_t v, v3<T> p)
{
    P q;
    q.x = p.x; q.y = p.y; q.z = p.z; q.w = 0;
    pos[v] = q;
}

// AbstractMeasurer::computeLocalElementSize (Measurers/AbstractMeasurer.cpp:137-155).
template<typename T, typename P> __device__ T localSize(const DevMesh& m, const P* pos, uint32_t v, v3<T> p)
{
    uint32_t b = m.noff[v], e = m.noff[v + 1];
    T total = T(0);
    for(uint32_t r = b; r < e; ++r)
    {
        v3<T> d = p - ld<T>(pos, m.nbr[r]);
        total += gmq::tsqrt(gmq::dot(d, d));
    }
    return total / (T)(e - b);
}

// FP32 reductions over aligned groups of G lanes (G a power of two): DPP
// moves inside rows of 16 lanes (lane ^ 1, lane ^ 2, then the mirrored half
// row and row, which pair lanes holding different partial results), a
// shuffle across rows.  The operation is commutative, so every lane of the
// group ends with the same bits.
template<int CTRL> __device__ __forceinline__ float dppMove(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF,
                                                                  false));
}

enum { DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140 };

template<int G, class Op> __device__ __forceinline__ float groupReduce(float v, Op op)
{
    if constexpr(G >= 2) v = op(v, dppMove<DPP_XOR1>(v));
    if constexpr(G >= 4) v = op(v, dppMove<DPP_XOR2>(v));
    if constexpr(G >= 8) v = op(v, dppMove<DPP_HALF_MIRROR>(v));
    if constexpr(G >= 16) v = op(v, dppMove<DPP_MIRROR>(v));
    if constexpr(G >= 32) v = op(v, __shfl_xor(v, 16, 64));
    if constexpr(G >= 64) v = op(v, __shfl_xor(v, 32, 64));
    return v;
}

template<int G> __device__ __forceinline__ float groupSum(float v)
{
    return groupReduce<G>(v, [](float a, float b) { return a + b; });
}

template<int G> __device__ __forceinline__ float groupMax(float v)
{
    return groupReduce<G>(v, [](float a, float b) { return __builtin_fmaxf(a, b); });
}

template<int G> __device__ __forceinline__ float groupMin(float v)
{
    return groupReduce<G>(v, [](float a, float b) { return __builtin_fminf(a, b); });
}

// FP64 group sum over shuffles (same tree, every lane gets the same bits).
template<int G> __device__ __forceinline__ double groupSum(double v)
{
#pragma unroll
    for(int off = G / 2; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

template<int G> __device__ __forceinline__ double groupMin(double v)
{
#pragma unroll
    for(int off = G / 2; off > 0; off >>= 1)
        v = __builtin_fmin(v, __shfl_xor(v, off, 64));
    return v;
}

// The local size with the neighbours spread over a group of W lanes (tree
// sum; FP32 with the hardware square root); every lane of the group gets the
// result.  Used by the FP32 kernels and the FP64 polish pass (not by the
// bit-exact FP64 mode, which sums in the reference's order).
template<int W>
__device__ __forceinline__ double localSizeGroup(const DevMesh& m, const double4* pos, uint32_t v, v3<double> p,
                                                 int lane)
{
    const uint32_t b = m.noff[v], e = m.noff[v + 1];
    double total = 0.0;
    for(uint32_t r = b + lane; r < e; r += W)
    {
        const v3<double> d = p - ld<double>(pos, m.nbr[r]);
        total += gmq::tsqrt(gmq::dot(d, d));
    }
    return groupSum<W>(total) / (double)(e - b);
}

template<int W>
__device__ __forceinline__ float localSizeGroup(const DevMesh& m, const float4* pos, uint32_t v, v3<float> p, int lane)
{
    const uint32_t b = m.noff[v], e = m.noff[v + 1];
    float total = 0.0f;
    for(uint32_t r = b + lane; r < e; r += W)
    {
        const v3<float> d = p - ld<float>(pos, m.nbr[r]);
        total += __builtin_amdgcn_sqrtf(gmq::dot(d, d));
    }
    return groupSum<W>(total) * __builtin_amdgcn_rcpf((float)(e - b));
}

// Records of the FP32 passes (mean ratio, metric-free): each element's
// quality in FP64 from the FP32 positions, with det(F)^(-2/3) from the
// hardware seed and two Newton steps (gma::powm23) and the reciprocal from
// the hardware estimate and two Newton steps, in place of the correctly
// rounded cube root and divisions of gmq::cornerQuality (a few ulp apart).
// Elements with a non-positive corner take the exact routine.  Returns q
// and sets inv = 1 / q for q > 0.
__device__ __forceinline__ double rcpNewton(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double evalCornerInv3(const gmq::m3<double>& F)
{
    const double d = gmq::det(F);
    const double f2 = gmq::dot(F.c[0], F.c[0]) + gmq::dot(F.c[1], F.c[1]) + gmq::dot(F.c[2], F.c[2]);
    return d > 0.0 ? f2 * gma::powm23(d) : __builtin_nan("");
}

// Cube root of x >= 0 for the mixed-precision FP64 passes: the hardware
// FP32 log/exp seed (about 2^-22 relative) and one Halley step
// y (y^3 + 2x) / (2 y^3 + x), whose cubic convergence leaves the last bits;
// the project's correctly rounded routine outside the FP32 range.
__device__ __forceinline__ double fastCbrt(double x)
{
    if(!(x >= 1.1754943508222875e-38 && x <= 3.4028234663852886e+38))
        return gm_cbrt_nonneg(x);
    const double y = (double)__builtin_amdgcn_exp2f(0.3333333333333333f * __builtin_amdgcn_logf((float)x));
    const double y3 = y * y * y;
    return y * (y3 + 2.0 * x) / (2.0 * y3 + x);
}

__device__ __forceinline__ double fastTetEval(const gmq::Crew& c, const v3<double> vp[4], double& inv)
{
    const double s = evalCornerInv3(
        gmq::mmul<double>(gmq::mcols<double>(vp[3] - vp[0], vp[3] - vp[1], vp[3] - vp[2]), gmq::frTetInv<double>()));
    if(s < __builtin_huge_val())
    {
        inv = s * (1.0 / 3.0);
        return 3.0 * rcpNewton(s);
    }
    const double q = gmq::tetQuality<double, 0, 0>(c, vp);
    inv = 1.0 / q;
    return q;
}

__device__ __forceinline__ double fastHexEval(const gmq::Crew& c, const v3<double> vp[8], double& inv)
{
    const v3<double> e01 = vp[1] - vp[0], e03 = vp[3] - vp[0], e04 = vp[4] - vp[0], e12 = vp[2] - vp[1];
    const v3<double> e15 = vp[5] - vp[1], e23 = vp[3] - vp[2], e26 = vp[6] - vp[2], e37 = vp[7] - vp[3];
    const v3<double> e45 = vp[5] - vp[4], e47 = vp[7] - vp[4], e56 = vp[6] - vp[5], e67 = vp[7] - vp[6];
    using gmq::mcols;
    const double s = evalCornerInv3(mcols<double>(e01, e04, -e03)) + evalCornerInv3(mcols<double>(e01, e12, e15)) +
                     evalCornerInv3(mcols<double>(e12, e26, -e23)) + evalCornerInv3(mcols<double>(e03, e23, e37)) +
                     evalCornerInv3(mcols<double>(e04, e45, e47)) + evalCornerInv3(mcols<double>(e15, -e56, e45)) +
                     evalCornerInv3(mcols<double>(e26, e56, e67)) + evalCornerInv3(mcols<double>(e37, e67, -e47));
    if(s < __builtin_huge_val())
    {
        // q = 8 / sum(1 / q_k) = 24 / sum(3 / q_k)
        inv = s * (1.0 / 24.0);
        return 24.0 * rcpNewton(s);
    }
    const double q = gmq::hexQuality<double, 0, 0>(c, vp);
    inv = 1.0 / q;
    return q;
}

// Constraint snapping of a sliding vertex (PlaneConstraint::operator(),
// Boundaries/Constraints/FaceConstraint.cpp:180-185; SegmentConstraint::
// operator(), EdgeConstraint.cpp:185-199), from its constraint table entry;
// the identity for a vertex without one (c = nullptr).
template<typename T> struct DevSnap
{
    const double* c;
    __device__ v3<T> operator()(v3<T> p) const
    {
        if(!c) return p;
        const v3<T> a = mk<T>((T)c[1], (T)c[2], (T)c[3]);
        const v3<T> n = mk<T>((T)c[4], (T)c[5], (T)c[6]);
        const v3<T> dist = p - a;
        const T l = gmq::dot(dist, n);
        if((int)c[0] == CONS_PLANE) return p - n * l;
        if(l < T(0)) return a;
        if(l > (T)c[10]) return mk<T>((T)c[7], (T)c[8], (T)c[9]);
        const v3<T> t = dist - n * l;
        return p - t;
    }
};

// Table entry of slot v's sliding constraint, nullptr if it has none.
__device__ __forceinline__ const double* consOf(const DevMesh& m, uint32_t v)
{
    if(!m.cons) return nullptr;
    const int32_t k = m.cons[v];
    return k >= 0 ? m.consTable + (size_t)CONS_STRIDE * k : nullptr;
}

// Quality of the patch element `ref` with the patch's centre vertex at p.
template<typename T, int EV, int SA, typename P>
__device__ __forceinline__ T elementQualityAt(const DevMesh& m, const gmq::Crew& c, const P* pos, uint32_t ref, v3<T> p)
{
    uint32_t id = ref >> 4;
    int k = (int)(ref & 7);
    if(!(ref & 8))
    {
        uint4 t = m.tets[id];
        v3<T> vp[4];
        vp[0] = k == 0 ? p : ld<T>(pos, t.x);
        vp[1] = k == 1 ? p : ld<T>(pos, t.y);
        vp[2] = k == 2 ? p : ld<T>(pos, t.z);
        vp[3] = k == 3 ? p : ld<T>(pos, t.w);
        return gmq::tetQuality<T, EV, SA>(c, vp);
    }
    uint4 a = m.hexs[2 * id], b = m.hexs[2 * id + 1];
    v3<T> vp[8];
    vp[0] = k == 0 ? p : ld<T>(pos, a.x);
    vp[1] = k == 1 ? p : ld<T>(pos, a.y);
    vp[2] = k == 2 ? p : ld<T>(pos, a.z);
    vp[3] = k == 3 ? p : ld<T>(pos, a.w);
    vp[4] = k == 4 ? p : ld<T>(pos, b.x);
    vp[5] = k == 5 ? p : ld<T>(pos, b.y);
    if(id >= m.nhexs)
        return gmq::priQuality<T, EV, SA>(c, vp);
    vp[6] = k == 6 ? p : ld<T>(pos, b.z);
    vp[7] = k == 7 ? p : ld<T>(pos, b.w);
    return gmq::hexQuality<T, EV, SA>(c, vp);
}

// AbstractEvaluator::patchQuality (Evaluators/AbstractEvaluator.cpp:236-279),
// one lane walking the whole patch.
template<typename T, int EV, int SA>
struct PlainPatchWalk
{
    const DevMesh& m;
    const gmq::Crew& c;
    const typename PosOf<T>::type* pos;
    uint32_t v;
    int count;

    __device__ __forceinline__ T at(v3<T> p)
    {
        ++count;
        T pq = T(0), pw = T(0);
        uint32_t b = m.eoff[v], e = m.eoff[v + 1];
        for(uint32_t r = b; r < e; ++r)
            gmq::accumulatePatch<T>(pq, pw, elementQualityAt<T, EV, SA>(m, c, pos, m.eref[r], p));
        return gmq::finalizePatch<T>(pq, pw);
    }
};

// The patch functor of the one-lane-per-vertex kernels; the compiler may
// keep it out of line.
template<typename T, int EV, int SA>
struct PlainPatch : PlainPatchWalk<T, EV, SA>
{
    __device__ T operator()(v3<T> p) { return this->at(p); }
};

// Under the texture sampler the walk is always inlined into its kernel.  Its
// out-of-line copy (metric conformity, FP64) is so long that the compiler
// expands its branches into s_getpc/s_setpc sequences through s[30:31], the
// register pair that holds the function's return address, without saving
// it: the function's final s_setpc_b64 s[30:31] then jumps back into its own
// body and the kernel never ends.  A kernel has no return address, so long
// branches inside it are harmless.  tests/test_isa_hazards.py checks every
// function of the built library for such an expansion.
template<typename T, int EV>
struct PlainPatch<T, EV, gmq::SA_TEXTURE> : PlainPatchWalk<T, EV, gmq::SA_TEXTURE>
{
    __device__ __forceinline__ T operator()(v3<T> p) { return this->at(p); }
};

// Plain kernel: one vertex of the work list per lane.
template<typename T, int EV, int SA, int SM>
__global__ void __launch_bounds__(256) smoothPlainKernel(const DevMesh* __restrict__ meshes,
                                                         const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                         gmq::Crew c, T moveCoeff)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count)
        return;
    typedef typename PosOf<T>::type P;
    const uint2 it = workItem(meshes, segs, nsegs, i);
    if(it.y == NO_ITEM || runStopped(meshes[it.x]))
        return;
    const DevMesh& m = meshes[it.x];
    P* pos = PosOf<T>::get(m);
    const uint32_t v = it.y;
    v3<T> p0 = ld<T>(pos, v);
    T ls = localSize<T>(m, pos, v, p0);
    PlainPatch<T, EV, SA> f{m, c, pos, v, 0};
    const DevSnap<T> snap{consOf(m, v)};
    v3<T> np;
    if(SM == SM_NELDER_MEAD) np = gms::nelderMead<T>(p0, ls, f, snap);
    else if(SM == SM_GRADIENT_DESCENT) np = gms::gradientDescent<T>(p0, ls, f, snap);
    else np = gms::spawnSearch<T>(p0, ls, moveCoeff, f, snap);
    st(pos, v, np);
    m.evalCounts[v] += f.count;
}

// ---------------------------------------------------------------------------
// Wavefront kernel: one vertex per group of G lanes (G a power of two >= the
// patch size), lane j of the group holding patch element j's other vertices
// in registers.  A patch evaluation computes the G element qualities in
// parallel and, in FP64, reduces them in patch order through LDS, so every
// lane of the group obtains the bit-identical sequential result and runs the
// optimiser in lock-step (FP32 reduces over a shuffle butterfly instead).
// 64 / G vertices share a wavefront; a block is one wavefront.
// ---------------------------------------------------------------------------
// EKIND: the patches' element kind (EK_TET or EK_HEX when the launch holds
// one kind only, EK_MIXED otherwise); a tetrahedron-only launch keeps four
// vertices per lane.
// FASTQ (FP64 mean ratio, metric-free; the polish pass): tetrahedron
// qualities from the hardware-seeded FP64 routines of the evaluation records
// instead of the correctly rounded ones.
template<typename T, int EV, int SA, int G, int EKIND = EK_MIXED, bool FASTQ = false>
struct WavePatch
{
    const gmq::Crew* c;
    T* red;           // LDS: 2 values per lane of this group (quality, 1/quality)
    int lane;         // lane within the group
    int nElem;        // patch size (<= G)
    int isHex;        // this lane's element is a record: a hexahedron, or a prism when isPri is set
    int isPri;
    int k;            // local index of the centre vertex in this lane's element
    v3<T> vp[EKIND == EK_TET ? 4 : 8];   // this lane's element (centre slot overwritten per call)
    int count;

    __device__ T operator()(v3<T> p)
    {
        ++count;
        return reduce(elemQuality(p));
    }

    // Quality of this lane's element with the centre vertex at p (1 past the
    // patch).
    __device__ T elemQuality(v3<T> p) const
    {
        T q = T(1);
        if(lane < nElem)
        {
            const bool hex = EKIND == EK_HEX || (EKIND == EK_MIXED && isHex);
            if(!hex)
            {
                v3<T> e[4];
#pragma unroll
                for(int i = 0; i < 4; ++i) e[i] = (i == k) ? p : vp[i];
                if constexpr(FASTQ && sizeof(T) == 8 && EV == 0 && SA == 0)
                {
                    double inv;
                    q = fastTetEval(*c, e, inv);
                }
                else
                    q = gmq::tetQuality<T, EV, SA>(*c, e);
            }
            else if constexpr(EKIND != EK_TET)
            {
                v3<T> e[8];
#pragma unroll
                for(int i = 0; i < 8; ++i) e[i] = (i == k) ? p : vp[i];
                if(EKIND == EK_MIXED && isPri)
                    q = gmq::priQuality<T, EV, SA>(*c, e);
                else
                    q = gmq::hexQuality<T, EV, SA>(*c, e);
            }
        }
        return q;
    }

    // Patch quality from the lanes' element qualities
    // (AbstractEvaluator::patchQuality rule); every lane of the group gets it.
    __device__ T reduce(T q)
    {
        if constexpr(std::is_same<T, float>::value)
        {
            // FP32: the same patch rule over a butterfly of the group's
            // lanes (tree order; deterministic, not the sequential sum).
            const bool valid = lane < nElem;
            const bool pos = valid && q > 0.0f;
            const float sum = groupSum<G>(pos ? __builtin_amdgcn_rcpf(q) : 0.0f);
            const float cnt = groupSum<G>(pos ? 1.0f : 0.0f);
            const bool bad = groupMax<G>((valid && !pos) ? 1.0f : 0.0f) > 0.0f;
            if(!bad)
                return cnt / sum;
            return groupMin<G>((valid && !pos && q == q) ? q : __builtin_huge_valf());
        }
        if constexpr(FASTQ)
        {
            // The mixed-precision FP64 passes: the same rule over the tree
            // (deterministic; the exact FP64 mode keeps the patch order).
            const bool valid = lane < nElem;
            const bool pos = valid && q > 0.0;
            const double sum = groupSum<G>(pos ? 1.0 / q : 0.0);
            const double cnt = groupSum<G>(pos ? 1.0 : 0.0);
            const bool bad = groupMin<G>((valid && !pos) ? 0.0 : 1.0) == 0.0;
            if(!bad)
                return cnt / sum;
            return groupMin<G>((valid && !pos && q == q) ? q : __builtin_huge_val());
        }
        T inv = T(1) / q;
        // LDS accesses of one wavefront complete in issue order, so the reads
        // below observe this group's writes; the barriers keep the compiler
        // from moving memory operations across the exchange.
        red[2 * lane] = q;
        red[2 * lane + 1] = inv;
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_wave_barrier();
        T pq = T(0), pw = T(0);
        for(int j = 0; j < nElem; ++j)
        {
            T qj = red[2 * j];
            if(pw != T(-1) && qj > T(0)) { pw += T(1); pq += red[2 * j + 1]; }
            else { pw = T(-1); pq = gmq::tmin(pq, qj); }
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_wave_barrier();
        return gmq::finalizePatch<T>(pq, pw);
    }
};

template<typename T, int EV, int SA, int SM, int G>
__global__ void __launch_bounds__(64) smoothWaveKernel(const DevMesh* __restrict__ meshes,
                                                       const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                       gmq::Crew c, T moveCoeff)
{
    typedef typename PosOf<T>::type P;
    __shared__ T red[2 * 64];
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const uint32_t i = blockIdx.x * (64 / G) + grp;
    // Groups past the end leave; groups never synchronise with each other.
    if(i >= count)
        return;
    const uint4 sp = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[sp.x];
    const uint32_t v = spanVertex(m, sp, i);
    if(v == NO_ITEM || runStopped(m))
        return;
    P* pos = PosOf<T>::get(m);
    WavePatch<T, EV, SA, G> f;
    f.c = &c;
    f.red = red + 2 * G * grp;
    f.lane = lane;
    f.count = 0;
    uint32_t eb = m.eoff[v], ee = m.eoff[v + 1];
    f.nElem = (int)(ee - eb);
    f.isHex = 0;
    f.isPri = 0;
    f.k = 0;
    if(lane < f.nElem)
    {
        uint32_t ref = m.eref[eb + lane];
        uint32_t id = ref >> 4;
        f.k = (int)(ref & 7);
        f.isHex = (ref >> 3) & 1;
        f.isPri = f.isHex && id >= m.nhexs;
        if(!f.isHex)
        {
            uint4 t = m.tets[id];
            f.vp[0] = ld<T>(pos, t.x); f.vp[1] = ld<T>(pos, t.y);
            f.vp[2] = ld<T>(pos, t.z); f.vp[3] = ld<T>(pos, t.w);
        }
        else
        {
            uint4 a = m.hexs[2 * id], b = m.hexs[2 * id + 1];
            f.vp[0] = ld<T>(pos, a.x); f.vp[1] = ld<T>(pos, a.y);
            f.vp[2] = ld<T>(pos, a.z); f.vp[3] = ld<T>(pos, a.w);
            f.vp[4] = ld<T>(pos, b.x); f.vp[5] = ld<T>(pos, b.y);
            f.vp[6] = ld<T>(pos, b.z); f.vp[7] = ld<T>(pos, b.w);
        }
    }
    v3<T> p0 = ld<T>(pos, v);
    T ls = localSize<T>(m, pos, v, p0);
    const DevSnap<T> snap{consOf(m, v)};
    v3<T> np;
    if(SM == SM_NELDER_MEAD) np = gms::nelderMead<T>(p0, ls, f, snap);
    else if(SM == SM_GRADIENT_DESCENT) np = gms::gradientDescent<T>(p0, ls, f, snap);
    else np = gms::spawnSearch<T>(p0, ls, moveCoeff, f, snap);
    if(lane == 0)
    {
        st(pos, v, np);
        m.evalCounts[v] += f.count;
    }
}

// ---------------------------------------------------------------------------
// Candidate-parallel kernels: the patch of a vertex is staged once in LDS
// (every element's vertex positions and the centre's local index), and each
// lane evaluates the whole patch for its own trial position, summing in patch
// order exactly like the sequential code.  Trial positions that the
// reference evaluates one after another but independently (the 64 spawn
// lattice points, the 6 gradient samples and the 8 line-search points) run
// side by side; the argmax keeps the reference's rule (strictly greater
// wins, so the first maximum; NaN never wins; all -inf/NaN -> index 0).
//   EK: 0 = tetrahedra only (4 slots per element), 1 = hexahedra only (8 slots).
// ---------------------------------------------------------------------------
template<int EK> struct Slots { static const int N = EK == 0 ? 4 : 8; };

template<typename T, int EK, typename P>
__device__ __forceinline__ void stageElement(const DevMesh& m, const P* pos, uint32_t ref, T* dst, uint8_t* k)
{
    uint32_t id = ref >> 4;
    *k = (uint8_t)(ref & 7);
    uint32_t vi[8];
    if(EK == 0)
    {
        uint4 t = m.tets[id];
        vi[0] = t.x; vi[1] = t.y; vi[2] = t.z; vi[3] = t.w;
    }
    else
    {
        uint4 a = m.hexs[2 * id], b = m.hexs[2 * id + 1];
        vi[0] = a.x; vi[1] = a.y; vi[2] = a.z; vi[3] = a.w;
        vi[4] = b.x; vi[5] = b.y; vi[6] = b.z; vi[7] = b.w;
    }
#pragma unroll
    for(int j = 0; j < Slots<EK>::N; ++j)
    {
        v3<T> q = ld<T>(pos, vi[j]);
        dst[3 * j] = q.x; dst[3 * j + 1] = q.y; dst[3 * j + 2] = q.z;
    }
}

// Patch quality (AbstractEvaluator::patchQuality order) from an LDS-staged patch.
template<typename T, int EV, int SA, int EK>
__device__ __forceinline__ T ldsPatchQuality(const gmq::Crew& c, const T* sp, const uint8_t* sk, int nE, v3<T> p)
{
    T pq = T(0), pw = T(0);
    for(int e = 0; e < nE; ++e)
    {
        const T* b = sp + e * Slots<EK>::N * 3;
        const int k = sk[e];
        T q;
        if(EK == 0)
        {
            v3<T> vp[4];
#pragma unroll
            for(int j = 0; j < 4; ++j) vp[j] = (j == k) ? p : mk<T>(b[3 * j], b[3 * j + 1], b[3 * j + 2]);
            q = gmq::tetQuality<T, EV, SA>(c, vp);
        }
        else
        {
            v3<T> vp[8];
#pragma unroll
            for(int j = 0; j < 8; ++j) vp[j] = (j == k) ? p : mk<T>(b[3 * j], b[3 * j + 1], b[3 * j + 2]);
            q = gmq::hexQuality<T, EV, SA>(c, vp);
        }
        gmq::accumulatePatch<T>(pq, pw, q);
    }
    return gmq::finalizePatch<T>(pq, pw);
}

// Argmax over `width` lanes (power of two, aligned groups): the largest key,
// lowest lane among equals.  NaN keys count as -inf.
template<typename T>
__device__ __forceinline__ int groupArgmax(T q, int lane, int width)
{
    T key = (q == q) ? q : (T)(-__builtin_huge_val());
    int idx = lane;
    for(int off = width >> 1; off > 0; off >>= 1)
    {
        T ok = __shfl_xor(key, off, 64);
        int oi = __shfl_xor(idx, off, 64);
        if(ok > key || (ok == key && oi < idx)) { key = ok; idx = oi; }
    }
    return idx;
}

// The same rule over an aligned group of G lanes: the largest key, the
// lowest lane among equal keys (a maximum reduction, then the first lane of
// the group holding it); returns the index within the group.  Every lane of
// the wavefront must be active.
template<int G> __device__ __forceinline__ int groupArgmaxFast(float key)
{
    key = (key == key) ? key : -__builtin_huge_valf();
    const float mx = groupMax<G>(key);
    const unsigned long long hit = __ballot(key == mx);
    const int first = (int)(threadIdx.x & 63) & ~(G - 1);
    const unsigned long long mine = G == 64 ? hit : (hit >> first) & ((1ull << G) - 1);
    return (int)__builtin_ctzll(mine);
}

template<typename T>
__device__ __forceinline__ v3<T> shflv(v3<T> a, int src)
{
    return mk<T>(__shfl(a.x, src, 64), __shfl(a.y, src, 64), __shfl(a.z, src, 64));
}

// LDS bytes of one wavefront's staging area: `elems` elements of kind EK.
template<typename T, int EK> __host__ __device__ inline size_t stageBytes(int elems)
{
    size_t b = (size_t)elems * Slots<EK>::N * 3 * sizeof(T);
    return b + (((size_t)elems + 15) & ~(size_t)15);
}

// Spawn search (SpawnSearchSmoother::smoothVertices, SpawnSearchSmoother.cpp:161-226):
// one vertex per wavefront, lane p evaluates lattice point p; `ppg`
// wavefronts (vertices) per block, each staging up to `cap` elements in LDS.
template<typename T, int EV, int SA, int EK>
__global__ void __launch_bounds__(512) spawnLaneKernel(const DevMesh* __restrict__ meshes,
                                                       const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                       gmq::Crew c, T moveCoeff, int cap)
{
    typedef typename PosOf<T>::type P;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ppg = blockDim.x >> 6;
    unsigned char* base = smem + (size_t)wave * stageBytes<T, EK>(cap);
    T* sp = (T*)base;
    uint8_t* sk = base + (size_t)cap * Slots<EK>::N * 3 * sizeof(T);
    const uint32_t i = xcdBlock() * ppg + wave;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    P* pos = PosOf<T>::get(m);
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    for(int e = lane; e < nE; e += 64)
        stageElement<T, EK>(m, pos, m.eref[eb + e], sp + e * Slots<EK>::N * 3, sk + e);
    __syncthreads();
    if(!present)
        return;
    v3<T> p0 = ld<T>(pos, v);
    T scale = moveCoeff * localSize<T>(m, pos, v, p0);
    for(int iter = 0; iter < 2; ++iter)
    {
        v3<T> prop = p0 + gms::spawnOffset<T>(lane) * scale;
        T q = ldsPatchQuality<T, EV, SA, EK>(c, sp, sk, nE, prop);
        int best = groupArgmax<T>(q, lane, 64);
        p0 = shflv<T>(prop, best);
        scale /= T(3);
    }
    if(lane == 0)
    {
        st(pos, v, p0);
        m.evalCounts[v] += 128;
    }
}

// Gradient descent (GradientDescentSmoother::smoothVertices,
// GradientDescentSmoother.cpp:73-202): 8 vertices per wavefront, 8 lanes each;
// lanes 0-5 evaluate the 6 gradient samples, lanes 0-7 the 8 line-search
// points.  `ppg` wavefronts per block; each vertex stages up to `cap` elements.
template<typename T, int EV, int SA, int EK>
__global__ void __launch_bounds__(512) gdLaneKernel(const DevMesh* __restrict__ meshes,
                                                    const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                    gmq::Crew c, int cap)
{
    typedef typename PosOf<T>::type P;
    extern __shared__ __align__(16) unsigned char smem[];
    const int SL = Slots<EK>::N * 3;
    const int lane = threadIdx.x;
    const int g = lane >> 3, s = lane & 7, base = lane & ~7;
    unsigned char* mine = smem + (size_t)g * stageBytes<T, EK>(cap);
    T* mySp = (T*)mine;
    uint8_t* mySk = mine + (size_t)cap * SL * sizeof(T);
    const uint32_t i = xcdBlock() * (blockDim.x >> 3) + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    P* pos = PosOf<T>::get(m);
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    for(int e = s; e < nE; e += 8)
        stageElement<T, EK>(m, pos, m.eref[eb + e], mySp + e * SL, mySk + e);
    __syncthreads();
    if(!present)
        return;

    v3<T> p = ld<T>(pos, v);
    T nodeShift = localSize<T>(m, pos, v, p) * T(1.0 / 75.0);
    const T originalNodeShift = nodeShift;
    int evals = 0;
    for(int cyc = 0; cyc < 5; ++cyc)
    {
        // Samples -x, +x, -y, +y, -z, +z on lanes 0..5.
        const T sx = s == 0 ? -nodeShift : (s == 1 ? nodeShift : T(0));
        const T sy = s == 2 ? -nodeShift : (s == 3 ? nodeShift : T(0));
        const T sz = s == 4 ? -nodeShift : (s == 5 ? nodeShift : T(0));
        T q = T(0);
        if(s < 6)
            q = ldsPatchQuality<T, EV, SA, EK>(c, mySp, mySk, nE, p + mk<T>(sx, sy, sz));
        evals += 6;
        const T q0 = __shfl(q, base + 0, 64), q1 = __shfl(q, base + 1, 64), q2 = __shfl(q, base + 2, 64);
        const T q3 = __shfl(q, base + 3, 64), q4 = __shfl(q, base + 4, 64), q5 = __shfl(q, base + 5, 64);
        v3<T> gq = mk<T>(q1 - q0, q3 - q2, q5 - q4);
        T gn = gmq::tsqrt(gmq::dot(gq, gq));
        if(gn == T(0))
            break;
        v3<T> shift = gq * (nodeShift / gn);
        const T off = s == 0 ? T(-0.25) : s == 1 ? T(0.00) : s == 2 ? T(0.10) : s == 3 ? T(0.20)
                    : s == 4 ? T(0.40) : s == 5 ? T(0.80) : s == 6 ? T(1.20) : T(1.60);
        v3<T> prop = p + shift * off;
        T qp = ldsPatchQuality<T, EV, SA, EK>(c, mySp, mySk, nE, prop);
        evals += 8;
        int best = groupArgmax<T>(qp, s, 8);
        p = shflv<T>(prop, (base & 63) + best);
        T bestOff = __shfl(off, (base & 63) + best, 64);
        nodeShift *= gmq::tabs(bestOff);
        if(nodeShift < originalNodeShift / T(10))
            break;
    }
    if(s == 0)
    {
        st(pos, v, p);
        m.evalCounts[v] += evals;
    }
}

// ---------------------------------------------------------------------------
// Mean ratio with the metric-free measurer in the mixed-precision mode:
// candidate-parallel kernels on affine corner terms (gm_affine.h) staged
// relative to the vertex's position, in FP32 for the bulk passes and in FP64
// for the polish pass.  Same searches as spawnLaneKernel / gdLaneKernel /
// Nelder-Mead; trial points are offsets x from the start position o, and the
// vertex moves to o + x.  (The exact FP64 mode keeps the kernels above, whose
// operation order matches the oracle bit for bit.)
// ---------------------------------------------------------------------------
template<int EV, int SA> struct AffineCrew { static const bool on = false; };
template<> struct AffineCrew<gmq::EV_MEAN_RATIO, gmq::SA_UNIFORM> { static const bool on = true; };

// True when the list runs on the affine kernels: every FP32 pass, and the
// FP64 polish pass, of the mean-ratio / metric-free crew.
template<typename T, int EV, int SA> __host__ __device__ inline bool useAffine(const VertexWork& w)
{
    return AffineCrew<EV, SA>::on && (sizeof(T) == 4 || w.polish);
}

template<typename T, int EK> __host__ __device__ inline size_t affineBytes()
{
    return EK == 0 ? sizeof(gma::Corner<T>) : sizeof(gma::Hex<T>);
}

// LDS bytes of `elems` staged elements of kind EK.
template<typename T, int EK> __host__ __device__ inline size_t laneStageBytes(int elems, bool affine)
{
    return affine ? (size_t)elems * affineBytes<T, EK>() : stageBytes<T, EK>(elems);
}

// FP64 determinant of a corner, from the FP32 positions with the moving
// vertex at p (widened exactly).
template<int EK>
struct DetFix
{
    const DevMesh* m;
    const float4* pos;

    __device__ v3<double> at(uint32_t v, int slot, int kk, v3<double> p) const
    {
        if(slot == kk) return p;
        const float4 q = pos[v];
        return mk<double>((double)q.x, (double)q.y, (double)q.z);
    }

    __device__ float det64(const gma::Corner<float>& k, v3<float> pf) const
    {
        const uint32_t ref = k.ref;
        const uint32_t id = ref >> 4;
        const int kk = (int)(ref & 7);
        const v3<double> p = mk<double>((double)pf.x, (double)pf.y, (double)pf.z);
        if(EK == 0)
        {
            const uint4 t = m->tets[id];
            const v3<double> v0 = at(t.x, 0, kk, p), v1 = at(t.y, 1, kk, p), v2 = at(t.z, 2, kk, p),
                             v3_ = at(t.w, 3, kk, p);
            const double dE = gmq::dot(v3_ - v0, gmq::cross(v3_ - v1, v3_ - v2));
            // det F = det E * det Fr_TET_INV, and det Fr_TET_INV = sqrt(2).
            return (float)(dE * 1.4142135623730951);
        }
        const uint4 a = m->hexs[2 * id], b = m->hexs[2 * id + 1];
        const uint32_t ids[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        const uint32_t code = gma::hexCornerCode((int)k.corner);
        v3<double> col[3];
#pragma unroll
        for(int j = 0; j < 3; ++j)
        {
            const int P = (int)((code >> (8 * j + 4)) & 15u), M = (int)((code >> (8 * j)) & 15u);
            uint32_t vp = ids[0], vm = ids[0];
#pragma unroll
            for(int q = 1; q < 8; ++q)
            {
                vp = P == q ? ids[q] : vp;
                vm = M == q ? ids[q] : vm;
            }
            col[j] = at(vp, P, kk, p) - at(vm, M, kk, p);
        }
        return (float)gmq::dot(col[0], gmq::cross(col[1], col[2]));
    }
};

// Constant corners of a staged hexahedron: the moving vertex at its position
// (FP32); FP64 staging needs no recheck.
template<typename T, int EK>
struct ConstFix
{
    const DevMesh* m;
    const float4* pos32;
    v3<T> o;
    __device__ float operator()(const gma::Corner<T>& k) const
    {
        if constexpr(sizeof(T) == 4)
            return DetFix<EK>{m, pos32}.det64(k, o);
        else
            return (float)k.d0;
    }
};

// FP64 inversion test of a chosen FP32 move (one staged element): returns
// INV_NONPOS32 when FP32 calls a corner non-positive, INV_FLIP when a corner
// FP32 calls positive lies inside the near-degenerate band and its FP64
// determinant (from the FP32 positions) is not positive.
enum { INV_NONPOS32 = 1, INV_FLIP = 2 };

template<int EK>
__device__ __forceinline__ int cornerInversion(const gma::Corner<float>& k, v3<float> x, v3<float> pf,
                                               const DetFix<EK>& fix)
{
    const float d = gma::cornerDet<float>(k, x);
    if(!(d > 0.0f)) return INV_NONPOS32;
    return (d < k.thr && !(fix.det64(k, pf) > 0.0f)) ? INV_FLIP : 0;
}

template<int EK>
__device__ __forceinline__ int inversionTest(const unsigned char* sp, int e, v3<float> x, v3<float> pf,
                                             const DetFix<EK>& fix)
{
    if(EK == 0)
        return cornerInversion<EK>(((const gma::Corner<float>*)sp)[e], x, pf, fix);
    const gma::Hex<float>& h = ((const gma::Hex<float>*)sp)[e];
    int flags = h.constMin <= 0.0f ? INV_NONPOS32 : 0;
    for(int c = 0; c < 4; ++c) flags |= cornerInversion<EK>(h.k[c], x, pf, fix);
    return flags;
}

// Applies the FP64 inversion test to a move of the vertex whose lanes are the
// `width` lanes starting at `first` (within the wavefront); returns the
// position to store.  FP64 moves need no test.
template<typename T, int EK>
__device__ __forceinline__ v3<T> testedMove(const DevMesh& m, const unsigned char* sp, int nE, int lane, int width,
                                            int first, v3<T> o, v3<T> x)
{
    const v3<T> pf = o + x;
    if constexpr(sizeof(T) == 8)
        return pf;
    else
    {
        const DetFix<EK> fix{&m, m.pos32};
        int flags = 0;
        for(int e = lane; e < nE; e += width)
            flags |= inversionTest<EK>(sp, e, x, pf, fix);
        const unsigned long long mask = width == 64 ? ~0ull : ((1ull << width) - 1);
        const bool reject = ((__ballot((flags & INV_FLIP) != 0) >> first) & mask) != 0 &&
                            ((__ballot((flags & INV_NONPOS32) != 0) >> first) & mask) == 0;
        return reject ? o : pf;
    }
}

template<typename T, int EK>
__device__ __forceinline__ void stageAffine(const DevMesh& m, const typename PosOf<T>::type* pos, uint32_t ref,
                                            v3<T> o, unsigned char* dst)
{
    const uint32_t id = ref >> 4;
    const int k = (int)(ref & 7);
    if(EK == 0)
    {
        const uint4 t = m.tets[id];
        v3<T> u[4] = {ld<T>(pos, t.x) - o, ld<T>(pos, t.y) - o, ld<T>(pos, t.z) - o, ld<T>(pos, t.w) - o};
        *(gma::Corner<T>*)dst = gma::tetCorner<T>(u, k, ref);
    }
    else
    {
        const uint4 a = m.hexs[2 * id], b = m.hexs[2 * id + 1];
        v3<T> u[8] = {ld<T>(pos, a.x) - o, ld<T>(pos, a.y) - o, ld<T>(pos, a.z) - o, ld<T>(pos, a.w) - o,
                      ld<T>(pos, b.x) - o, ld<T>(pos, b.y) - o, ld<T>(pos, b.z) - o, ld<T>(pos, b.w) - o};
        gma::stageHex<T>(*(gma::Hex<T>*)dst, u, k, ref, ConstFix<T, 1>{&m, m.pos32, o});
    }
}

// Sum of 3/q over the patch assuming every element is positive
// (branch-free; not finite when the assumption fails).
template<typename T, int EK>
__device__ __forceinline__ T affineSumFast(const unsigned char* sp, int nE, v3<T> x, T xx)
{
    T inv3 = T(0);
    for(int e = 0; e < nE; ++e)
    {
        if(EK == 0) inv3 += gma::fastTet<T>(((const gma::Corner<T>*)sp)[e], x, xx);
        else inv3 += gma::fastHex<T>(((const gma::Hex<T>*)sp)[e], x, xx);
    }
    return inv3;
}

// affineSumFast at P offsets, each staged row read once.
template<typename T, int EK, int P>
__device__ __forceinline__ void affineSumsFast(const unsigned char* sp, int nE, const v3<T> (&x)[P], const T (&xx)[P],
                                               T (&inv3)[P])
{
#pragma unroll
    for(int p = 0; p < P; ++p) inv3[p] = T(0);
    for(int e = 0; e < nE; ++e)
    {
        if(EK == 0)
        {
            const gma::Corner<T> k = ((const gma::Corner<T>*)sp)[e];
#pragma unroll
            for(int p = 0; p < P; ++p) inv3[p] += gma::fastTet<T>(k, x[p], xx[p]);
        }
        else
        {
            const gma::Hex<T>& h = ((const gma::Hex<T>*)sp)[e];
#pragma unroll
            for(int p = 0; p < P; ++p) inv3[p] += gma::fastHex<T>(h, x[p], xx[p]);
        }
    }
}

// Patch quality assuming every element is positive (branch-free); `bad`
// reports that the assumption failed.
template<typename T, int EK>
__device__ __forceinline__ T affinePatchFast(const unsigned char* sp, int nE, v3<T> x, T xx, bool& bad)
{
    const T inv3 = affineSumFast<T, EK>(sp, nE, x, xx);
    bad = gma::fastBad<T>(inv3);
    if constexpr(sizeof(T) == 4)
        return 3.0f * (float)nE * __builtin_amdgcn_rcpf(inv3);
    else
        return T(3) * (T)nE / inv3;
}

// The exact patch rule (smallest non-positive quality when one exists).
template<typename T, int EK>
__device__ __forceinline__ T affinePatchExact(const unsigned char* sp, int nE, v3<T> x, T xx)
{
    gma::Acc<T> a;
    for(int e = 0; e < nE; ++e)
    {
        if(EK == 0) gma::addTet<T>(a, ((const gma::Corner<T>*)sp)[e], x, xx);
        else gma::addHex<T>(a, ((const gma::Hex<T>*)sp)[e], x, xx);
    }
    return a.finish();
}

template<typename T, int EK>
__device__ __forceinline__ T affinePatch(const unsigned char* sp, int nE, v3<T> x)
{
    const T xx = gmq::dot(x, x);
    bool bad = false;
    const T q = affinePatchFast<T, EK>(sp, nE, x, xx, bad);
    return bad ? affinePatchExact<T, EK>(sp, nE, x, xx) : q;
}

// A value of lane `src` (wavefront-uniform; every lane must be active).
__device__ __forceinline__ float readLane(float v, int src)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}

__device__ __forceinline__ double readLane(double v, int src)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Stages the affine terms of the patches of the VPW vertices of a wavefront
// (vertex h's lanes h L .. h L + L - 1, L = 64 / VPW) into their tiles
// waveSp + h tileBytes, all 64 lanes sharing the rows: row r of the
// wavefront is row r - start(h) of vertex h.  Vertex h's element range and
// position are read from its lane h L while every lane is active (a
// cross-lane read from a lane that is off returns zero, and the staging loop
// runs on fewer lanes).  Returns the wavefront's largest patch.
template<typename T, int EK, int VPW>
__device__ __forceinline__ int stageWaveRows(const DevMesh& m, const typename PosOf<T>::type* pos,
                                             unsigned char* waveSp, size_t tileBytes, int nE, uint32_t eb, v3<T> o)
{
    constexpr int L = 64 / VPW;
    const int lane = threadIdx.x & 63;
    int start[VPW + 1];
    uint32_t ebOf[VPW];
    v3<T> oOf[VPW];
    start[0] = 0;
#pragma unroll
    for(int h = 0; h < VPW; ++h)
    {
        start[h + 1] = start[h] + __builtin_amdgcn_readlane(nE, h * L);
        ebOf[h] = (uint32_t)__builtin_amdgcn_readlane((int)eb, h * L);
        oOf[h] = mk<T>(readLane(o.x, h * L), readLane(o.y, h * L), readLane(o.z, h * L));
    }
    int nEmax = 0;
#pragma unroll
    for(int h = 0; h < VPW; ++h) nEmax = max(nEmax, start[h + 1] - start[h]);
    for(int r = lane; r < start[VPW]; r += 64)
    {
        int h = 0;
        uint32_t ebH = ebOf[0];
        v3<T> oH = oOf[0];
#pragma unroll
        for(int t = 1; t < VPW; ++t)
            if(r >= start[t])
            {
                h = t;
                ebH = ebOf[t];
                oH = oOf[t];
            }
        const int e = r - start[h];
        stageAffine<T, EK>(m, pos, m.eref[ebH + e], oH, waveSp + (size_t)h * tileBytes + (size_t)e * affineBytes<T, EK>());
    }
    return nEmax;
}

// Largest value over aligned groups of G lanes (FP64, shuffles).
template<int G> __device__ __forceinline__ double groupMax(double v)
{
#pragma unroll
    for(int off = G / 2; off > 0; off >>= 1)
        v = __builtin_fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// The first lattice point, in the order j + L p of a vertex's lanes j and
// sweep steps p, whose key equals the group's largest; `mask` selects the
// vertex's lanes, starting at lane `first`.
template<typename T, int P, int L>
__device__ __forceinline__ int firstBest(const T (&key)[P], T mx, unsigned long long mask, int first)
{
    int cand = 0;
#pragma unroll
    for(int p = P - 1; p >= 0; --p)
    {
        const unsigned long long hit = (__ballot(key[p] == mx) & mask) >> first;
        if(hit) cand = (int)__builtin_ctzll(hit) + L * p;
    }
    return cand;
}

// ---------------------------------------------------------------------------
// Matrix-core evaluation of FP32 affine corner terms.  A patch's corner rows
// (a tetrahedron's corner, or the four corners of a hexahedron that move
// with the vertex) go four at a time through v_mfma_f32_4x4x1f32 (16 blocks
// of 4x4, K = 1): lane l supplies the coefficient of row 4g + (l & 3) as the
// A operand and the coordinate of its own candidate as B, and receives
// D[r] = A(lane 4 (l / 4) + r) * B(lane l) + C[r] -- rows 4g..4g+3 at its
// candidate -- in every 4-lane block.  Two chains build
//   d  = d0 + gx x + gy y + gz z,
//   f2 = c0 + hx x + hy y + hz z + s |x|^2,
// and the vector ALUs only form f2 |d|^(-2/3) and the sum.  Every lane of the
// wavefront takes part (the matrix instructions run on the whole wavefront);
// the row count is a multiple of 4 (tetrahedron patches are padded with
// neutral rows, d = 1 and f2 = 0).
// ---------------------------------------------------------------------------
typedef float mfv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfv4 mfma4x4(float a, float b, mfv4 c)
{
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

template<int EK>
__device__ __forceinline__ const gma::Corner<float>& cornerRow(const unsigned char* sp, int r)
{
    if(EK == 0) return ((const gma::Corner<float>*)sp)[r];
    return ((const gma::Hex<float>*)sp)[r >> 2].k[r & 3];
}

// Sum over the patch's corner rows of f2 |d|^(-2/3) at this lane's candidate
// offset x (|x|^2 = xx); NaN or +inf when a corner is not positive.
template<int EK>
__device__ __forceinline__ float mfmaRowSum(const unsigned char* sp, int nRows, v3<float> x, float xx)
{
    const int sub = threadIdx.x & 3;
    float sum = 0.0f;
    for(int r0 = 0; r0 < nRows; r0 += 4)
    {
        const gma::Corner<float>& k = cornerRow<EK>(sp, r0 + sub);
        mfv4 d = {0.0f, 0.0f, 0.0f, 0.0f}, f = {0.0f, 0.0f, 0.0f, 0.0f};
        d = mfma4x4(k.d0, 1.0f, d);
        f = mfma4x4(k.c0, 1.0f, f);
        d = mfma4x4(k.gx, x.x, d);
        f = mfma4x4(k.hx, x.x, f);
        d = mfma4x4(k.gy, x.y, d);
        f = mfma4x4(k.hy, x.y, f);
        d = mfma4x4(k.gz, x.z, d);
        f = mfma4x4(k.hz, x.z, f);
        f = mfma4x4(k.s, xx, f);
#pragma unroll
        for(int r = 0; r < 4; ++r) sum += f[r] * gma::powm23(d[r]);
    }
    return sum;
}

// The sums of mfmaRowSum at P candidates per lane, each row's coefficients
// read from LDS once; the constant terms (d0, c0) go through the matrix
// cores once per row block and start every candidate's chains.
template<int EK, int P>
__device__ __forceinline__ void mfmaRowSums(const unsigned char* sp, int nRows, const v3<float> (&x)[P],
                                            const float (&xx)[P], float (&sum)[P])
{
    const int sub = threadIdx.x & 3;
#pragma unroll
    for(int p = 0; p < P; ++p) sum[p] = 0.0f;
    for(int r0 = 0; r0 < nRows; r0 += 4)
    {
        const gma::Corner<float>& k = cornerRow<EK>(sp, r0 + sub);
        const float gx = k.gx, gy = k.gy, gz = k.gz, hx = k.hx, hy = k.hy, hz = k.hz, ks = k.s;
        const mfv4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
        const mfv4 d0 = mfma4x4(k.d0, 1.0f, zero);
        const mfv4 c0 = mfma4x4(k.c0, 1.0f, zero);
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            mfv4 d = mfma4x4(gx, x[p].x, d0);
            mfv4 f = mfma4x4(hx, x[p].x, c0);
            d = mfma4x4(gy, x[p].y, d);
            f = mfma4x4(hy, x[p].y, f);
            d = mfma4x4(gz, x[p].z, d);
            f = mfma4x4(hz, x[p].z, f);
            f = mfma4x4(ks, xx[p], f);
#pragma unroll
            for(int r = 0; r < 4; ++r) sum[p] += f[r] * gma::powm23(d[r]);
        }
    }
}

// Vertices per wavefront of the FP32 spawn-search kernel.
constexpr int SPAWN_VPW = 4;

// Spawn search, FP32 affine terms on the matrix cores (the schedule, the
// candidates and the rules of spawnAffineKernel), VPW vertices per
// wavefront.  The L = 64 / VPW lanes of vertex g (lanes g L ...) take its
// lattice points j, j + L, ... (P = VPW per lane); a matrix instruction's
// 4-lane blocks each read their own vertex's rows, so one instruction
// serves every vertex of the wavefront.  All lanes of the wavefront stage
// the rows of all its vertices together, and every vertex's rows are padded
// with neutral ones (d = 1, f2 = 0: they add an exact zero) to the
// wavefront's largest row count, so a vertex's result does not depend on
// its neighbours in the wavefront.  The staging, the local size, the argmax
// and the inversion test are shared by VPW vertices instead of one.
template<int EK, int VPW>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(EK == 0 ? 8 : 1))) spawnMfmaKernel(const DevMesh* __restrict__ meshes,
                                                       const uint4* __restrict__ segs, uint32_t nsegs,
                                                       uint32_t count, float moveCoeff, int cap)
{
    constexpr int L = 64 / VPW;
    constexpr int P = VPW;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ppg = blockDim.x >> 6;
    const int g = lane / L, j = lane % L;
    const size_t tileBytes = (size_t)cap * affineBytes<float, EK>();
    unsigned char* waveSp = smem + (size_t)wave * VPW * tileBytes;
    unsigned char* sp = waveSp + (size_t)g * tileBytes;
    const uint32_t i = (xcdBlock() * ppg + wave) * VPW + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    float4* pos = m.pos32;
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<float> o = present ? ld<float>(pos, v) : mk<float>(0.0f, 0.0f, 0.0f);
    const int nEmax = stageWaveRows<float, EK, VPW>(m, pos, waveSp, tileBytes, nE, eb, o);
    // Neutral rows (tetrahedra: up to a multiple of 4 rows; hexahedra: whole
    // neutral elements) up to the wavefront's largest patch.
    const int nPad = EK == 0 ? (nEmax + 3) & ~3 : nEmax;
    for(int e = nE + j; e < nPad; e += L)
    {
        if(EK == 0)
        {
            gma::Corner<float> z{};
            z.d0 = 1.0f;
            ((gma::Corner<float>*)sp)[e] = z;
        }
        else
        {
            gma::Hex<float> z{};
            for(int c = 0; c < 4; ++c) z.k[c].d0 = 1.0f;
            z.constInv3 = 0.0f;
            z.constMin = 1.0f;
            ((gma::Hex<float>*)sp)[e] = z;
        }
    }
    __syncthreads();
    if(__ballot(present) == 0)
        return;
    const int nRows = EK == 0 ? nPad : 4 * nPad;
    float constSum = 0.0f;
    if(EK == 1)
        for(int e = 0; e < nE; ++e) constSum += ((const gma::Hex<float>*)sp)[e].constInv3;
    float scale = present ? moveCoeff * localSizeGroup<L>(m, pos, v, o, j) : 0.0f;
    const unsigned long long groupMask = (L == 64 ? ~0ull : ((1ull << L) - 1)) << (g * L);
    v3<float> xb = mk<float>(0.0f, 0.0f, 0.0f);
    for(int iter = 0; iter < 2; ++iter)
    {
        v3<float> x[P];
        float xx[P], sum[P];
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            x[p] = xb + gms::spawnOffset<float>(j + L * p) * scale;
            xx[p] = gmq::dot(x[p], x[p]);
        }
        mfmaRowSums<EK, P>(sp, nRows, x, xx, sum);
        // The patch quality 3 n / sum falls as the sum grows, so the best
        // candidate has the smallest sum.  As in spawnAffineKernel, a
        // candidate with a non-positive element loses unless every candidate
        // of the vertex has one; then the exact rule decides.
        bool anyGood = false;
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            if(EK == 1) sum[p] = 0.125f * (sum[p] + constSum);
            anyGood |= !gma::fastBad<float>(sum[p]);
        }
        const bool exact = (__ballot(anyGood) & groupMask) == 0;
        float key[P];
        float best = -__builtin_huge_valf();
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            if(exact)
                key[p] = affinePatchExact<float, EK>(sp, nE, x[p], xx[p]);
            else
                key[p] = gma::fastBad<float>(sum[p]) ? -__builtin_huge_valf() : -sum[p];
            key[p] = key[p] == key[p] ? key[p] : -__builtin_huge_valf();
            best = __builtin_fmaxf(best, key[p]);
        }
        const int cand = firstBest<float, P, L>(key, groupMax<L>(best), groupMask, g * L);
        xb = xb + gms::spawnOffset<float>(cand) * scale;
        scale *= 1.0f / 3.0f;
    }
    if(!present)
        return;
    const v3<float> np = testedMove<float, EK>(m, sp, nE, j, L, g * L, o, xb);
    if(j == 0)
    {
        st(pos, v, np);
        m.evalCounts[v] += 128;
    }
}

// Spawn search in FP64 on the vector ALUs (the polish pass), with the
// layout of spawnMfmaKernel: VPW vertices per wavefront, L = 64 / VPW lanes
// each taking lattice points j, j + L, ...; each lane sums its vertex's
// staged rows in element order.
template<int EK, int VPW>
__global__ void __launch_bounds__(512) spawnAffineKernel(const DevMesh* __restrict__ meshes,
                                                         const uint4* __restrict__ segs, uint32_t nsegs,
                                                         uint32_t count, double moveCoeff, int cap)
{
    constexpr int L = 64 / VPW;
    constexpr int P = VPW;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ppg = blockDim.x >> 6;
    const int g = lane / L, j = lane % L;
    const size_t tileBytes = (size_t)cap * affineBytes<double, EK>();
    unsigned char* waveSp = smem + (size_t)wave * VPW * tileBytes;
    const unsigned char* sp = waveSp + (size_t)g * tileBytes;
    const uint32_t i = (xcdBlock() * ppg + wave) * VPW + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    double4* pos = m.pos64;
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<double> o = present ? ld<double>(pos, v) : mk<double>(0.0, 0.0, 0.0);
    stageWaveRows<double, EK, VPW>(m, pos, waveSp, tileBytes, nE, eb, o);
    __syncthreads();
    if(__ballot(present) == 0)
        return;
    double scale = present ? moveCoeff * localSizeGroup<L>(m, pos, v, o, j) : 0.0;
    const unsigned long long groupMask = (L == 64 ? ~0ull : ((1ull << L) - 1)) << (g * L);
    v3<double> xb = mk<double>(0.0, 0.0, 0.0);
    for(int iter = 0; iter < 2; ++iter)
    {
        // The best candidate has the smallest sum of 3/q.  A candidate with a
        // non-positive element has quality <= 0 and loses to any candidate
        // without one; its exact value matters only when every candidate of
        // the vertex has one.
        v3<double> x[P];
        double xx[P], sum[P];
        bool anyGood = false;
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            x[p] = xb + gms::spawnOffset<double>(j + L * p) * scale;
            xx[p] = gmq::dot(x[p], x[p]);
        }
        affineSumsFast<double, EK, P>(sp, nE, x, xx, sum);
#pragma unroll
        for(int p = 0; p < P; ++p) anyGood |= !gma::fastBad<double>(sum[p]);
        const bool exact = (__ballot(anyGood) & groupMask) == 0;
        double key[P];
        double best = -__builtin_huge_val();
#pragma unroll
        for(int p = 0; p < P; ++p)
        {
            if(exact)
                key[p] = affinePatchExact<double, EK>(sp, nE, x[p], xx[p]);
            else
                key[p] = gma::fastBad<double>(sum[p]) ? -__builtin_huge_val() : -sum[p];
            key[p] = key[p] == key[p] ? key[p] : -__builtin_huge_val();
            best = __builtin_fmax(best, key[p]);
        }
        const int cand = firstBest<double, P, L>(key, groupMax<L>(best), groupMask, g * L);
        xb = xb + gms::spawnOffset<double>(cand) * scale;
        scale /= 3.0;
    }
    if(!present)
        return;
    if(j == 0)
    {
        st(pos, v, o + xb);
        m.evalCounts[v] += 128;
    }
}

template<typename T, int EK>
__global__ void __launch_bounds__(512) gdAffineKernel(const DevMesh* __restrict__ meshes,
                                                      const uint4* __restrict__ segs, uint32_t nsegs,
                                                      uint32_t count, int cap)
{
    typedef typename PosOf<T>::type P;
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int g = lane >> 3, s = lane & 7, base = lane & ~7;
    unsigned char* mySp = smem + (size_t)g * cap * affineBytes<T, EK>();
    const uint32_t i = xcdBlock() * (blockDim.x >> 3) + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    P* pos = PosOf<T>::get(m);
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<T> o = present ? ld<T>(pos, v) : mk<T>(T(0), T(0), T(0));
    for(int e = s; e < nE; e += 8)
        stageAffine<T, EK>(m, pos, m.eref[eb + e], o, mySp + (size_t)e * affineBytes<T, EK>());
    __syncthreads();
    if(!present)
        return;

    v3<T> x = mk<T>(T(0), T(0), T(0));
    T nodeShift = localSizeGroup<8>(m, pos, v, o, s) * T(1.0 / 75.0);
    const T originalNodeShift = nodeShift;
    int evals = 0;
    for(int cyc = 0; cyc < 5; ++cyc)
    {
        const T sx = s == 0 ? -nodeShift : (s == 1 ? nodeShift : T(0));
        const T sy = s == 2 ? -nodeShift : (s == 3 ? nodeShift : T(0));
        const T sz = s == 4 ? -nodeShift : (s == 5 ? nodeShift : T(0));
        T q = T(0);
        if(s < 6)
            q = affinePatch<T, EK>(mySp, nE, x + mk<T>(sx, sy, sz));
        evals += 6;
        const T q0 = __shfl(q, base + 0, 64), q1 = __shfl(q, base + 1, 64), q2 = __shfl(q, base + 2, 64);
        const T q3 = __shfl(q, base + 3, 64), q4 = __shfl(q, base + 4, 64), q5 = __shfl(q, base + 5, 64);
        const v3<T> gq = mk<T>(q1 - q0, q3 - q2, q5 - q4);
        const T gn = gmq::tsqrt(gmq::dot(gq, gq));
        if(gn == T(0))
            break;
        const v3<T> shift = gq * (nodeShift / gn);
        const T off = s == 0 ? T(-0.25) : s == 1 ? T(0.00) : s == 2 ? T(0.10) : s == 3 ? T(0.20)
                    : s == 4 ? T(0.40) : s == 5 ? T(0.80) : s == 6 ? T(1.20) : T(1.60);
        const v3<T> prop = x + shift * off;
        const T qp = affinePatch<T, EK>(mySp, nE, prop);
        evals += 8;
        const int best = groupArgmax<T>(qp, s, 8);
        x = shflv<T>(prop, (base & 63) + best);
        const T bestOff = __shfl(off, (base & 63) + best, 64);
        nodeShift *= gmq::tabs(bestOff);
        if(nodeShift < originalNodeShift / T(10))
            break;
    }
    const v3<T> np = testedMove<T, EK>(m, mySp, nE, s, 8, (threadIdx.x & 63) & ~7, o, x);
    if(s == 0)
    {
        st(pos, v, np);
        m.evalCounts[v] += evals;
    }
}

// Gradient descent in FP64 on tetrahedral patches (the polish pass): the
// schedule of gdAffineKernel with 72-byte LDS rows (the nine terms a
// candidate reads; FP64 moves need no inversion test, so nothing else).
struct RowCoefD
{
    double d0, gx, gy, gz, c0, hx, hy, hz, s;
};

static __device__ __forceinline__ double coefPatchQuality(const RowCoefD* rows, int nE, v3<double> x)
{
    const double xx = gmq::dot(x, x);
    double inv3 = 0.0;
    for(int e = 0; e < nE; ++e)
    {
        const RowCoefD& r = rows[e];
        gma::Corner<double> c{};
        c.d0 = r.d0; c.gx = r.gx; c.gy = r.gy; c.gz = r.gz;
        c.c0 = r.c0; c.hx = r.hx; c.hy = r.hy; c.hz = r.hz; c.s = r.s;
        inv3 += gma::fastTet<double>(c, x, xx);
    }
    if(!gma::fastBad<double>(inv3))
        return 3.0 * (double)nE / inv3;
    gma::Acc<double> a;
    for(int e = 0; e < nE; ++e)
    {
        const RowCoefD& r = rows[e];
        gma::Corner<double> c{};
        c.d0 = r.d0; c.gx = r.gx; c.gy = r.gy; c.gz = r.gz;
        c.c0 = r.c0; c.hx = r.hx; c.hy = r.hy; c.hz = r.hz; c.s = r.s;
        gma::addTet<double>(a, c, x, xx);
    }
    return a.finish();
}

static __global__ void __launch_bounds__(512) gdAffineTetKernel(const DevMesh* __restrict__ meshes,
                                                         const uint4* __restrict__ segs, uint32_t nsegs,
                                                         uint32_t count, int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int g = lane >> 3, s = lane & 7, base = lane & ~7;
    RowCoefD* rows = (RowCoefD*)smem + (size_t)g * cap;
    const uint32_t i = xcdBlock() * (blockDim.x >> 3) + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    double4* pos = m.pos64;
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<double> o = present ? ld<double>(pos, v) : mk<double>(0.0, 0.0, 0.0);
    for(int e = s; e < nE; e += 8)
    {
        const uint32_t ref = m.eref[eb + e];
        const uint4 t = m.tets[ref >> 4];
        v3<double> u[4] = {ld<double>(pos, t.x) - o, ld<double>(pos, t.y) - o, ld<double>(pos, t.z) - o,
                           ld<double>(pos, t.w) - o};
        const gma::Corner<double> c = gma::tetCorner<double>(u, (int)(ref & 7), ref);
        rows[e] = RowCoefD{c.d0, c.gx, c.gy, c.gz, c.c0, c.hx, c.hy, c.hz, c.s};
    }
    __syncthreads();
    if(!present)
        return;

    v3<double> x = mk<double>(0.0, 0.0, 0.0);
    double nodeShift = localSizeGroup<8>(m, pos, v, o, s) * (1.0 / 75.0);
    const double originalNodeShift = nodeShift;
    int evals = 0;
    for(int cyc = 0; cyc < 5; ++cyc)
    {
        const double sx = s == 0 ? -nodeShift : (s == 1 ? nodeShift : 0.0);
        const double sy = s == 2 ? -nodeShift : (s == 3 ? nodeShift : 0.0);
        const double sz = s == 4 ? -nodeShift : (s == 5 ? nodeShift : 0.0);
        double q = 0.0;
        if(s < 6)
            q = coefPatchQuality(rows, nE, x + mk<double>(sx, sy, sz));
        evals += 6;
        const double q0 = __shfl(q, base + 0, 64), q1 = __shfl(q, base + 1, 64), q2 = __shfl(q, base + 2, 64);
        const double q3 = __shfl(q, base + 3, 64), q4 = __shfl(q, base + 4, 64), q5 = __shfl(q, base + 5, 64);
        const v3<double> gq = mk<double>(q1 - q0, q3 - q2, q5 - q4);
        const double gn = gmq::tsqrt(gmq::dot(gq, gq));
        if(gn == 0.0)
            break;
        const v3<double> shift = gq * (nodeShift / gn);
        const double off = s == 0 ? -0.25 : s == 1 ? 0.00 : s == 2 ? 0.10 : s == 3 ? 0.20
                         : s == 4 ? 0.40 : s == 5 ? 0.80 : s == 6 ? 1.20 : 1.60;
        const v3<double> prop = x + shift * off;
        const double qp = coefPatchQuality(rows, nE, prop);
        evals += 8;
        const int best = groupArgmax<double>(qp, s, 8);
        x = shflv<double>(prop, (base & 63) + best);
        const double bestOff = __shfl(off, (base & 63) + best, 64);
        nodeShift *= gmq::tabs(bestOff);
        if(nodeShift < originalNodeShift / 10.0)
            break;
    }
    if(s == 0)
    {
        st(pos, v, o + x);
        m.evalCounts[v] += evals;
    }
}

// Patch quality of this lane's candidate from the FP32 affine rows on the
// matrix cores (every lane of the wavefront takes part); the exact rule, on
// the vector ALUs, for a lane whose patch has a non-positive element.
template<int EK>
__device__ __forceinline__ float mfmaPatchQuality(const unsigned char* sp, int nRows, int nE, v3<float> x,
                                                  float constSum)
{
    const float xx = gmq::dot(x, x);
    float sum = mfmaRowSum<EK>(sp, nRows, x, xx);
    if(EK == 1) sum = 0.125f * (sum + constSum);
    if(!gma::fastBad<float>(sum))
        return 3.0f * (float)nE * __builtin_amdgcn_rcpf(sum);
    return affinePatchExact<float, EK>(sp, nE, x, xx);
}

// Gradient descent, FP32 affine terms on the matrix cores: the schedule,
// samples and proposals of gdAffineKernel (8 vertices per wavefront, 8 lanes
// each).  The matrix instructions need the whole wavefront, so the cycle loop
// runs while any of its vertices is still descending; a vertex that stopped
// keeps its position, step and count.  Each vertex's rows are padded with
// neutral ones up to the wavefront's largest patch.
template<int EK>
__global__ void __launch_bounds__(512) gdMfmaKernel(const DevMesh* __restrict__ meshes,
                                                    const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                    int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int g = lane >> 3, s = lane & 7, base = lane & ~7;
    unsigned char* mySp = smem + (size_t)g * cap * affineBytes<float, EK>();
    const uint32_t i = xcdBlock() * (blockDim.x >> 3) + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    float4* pos = m.pos32;
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<float> o = present ? ld<float>(pos, v) : mk<float>(0.0f, 0.0f, 0.0f);
    for(int e = s; e < nE; e += 8)
        stageAffine<float, EK>(m, pos, m.eref[eb + e], o, mySp + (size_t)e * affineBytes<float, EK>());
    // Rows up to the wavefront's largest patch (rounded to 4 rows); the tile
    // holds them (a multiple of 4 elements, at least the largest patch).
    int nEmax = nE;
#pragma unroll
    for(int off = 8; off < 64; off <<= 1)
        nEmax = max(nEmax, __shfl_xor(nEmax, off, 64));
    const int nPad = EK == 0 ? (nEmax + 3) & ~3 : nEmax;
    for(int e = nE + s; e < nPad; e += 8)
    {
        if(EK == 0)
        {
            gma::Corner<float> z{};
            z.d0 = 1.0f;
            ((gma::Corner<float>*)mySp)[e] = z;
        }
        else
        {
            gma::Hex<float> z{};
            for(int c = 0; c < 4; ++c) z.k[c].d0 = 1.0f;
            z.constInv3 = 0.0f;
            z.constMin = 1.0f;
            ((gma::Hex<float>*)mySp)[e] = z;
        }
    }
    __syncthreads();
    const int nRows = EK == 0 ? nPad : 4 * nPad;
    float constSum = 0.0f;
    if(EK == 1)
        for(int e = 0; e < nE; ++e) constSum += ((const gma::Hex<float>*)mySp)[e].constInv3;

    v3<float> x = mk<float>(0.0f, 0.0f, 0.0f);
    float nodeShift = localSizeGroup<8>(m, pos, v, o, s) * (1.0f / 75.0f);
    const float originalNodeShift = nodeShift;
    int evals = 0;
    bool active = present;
    const float off = s == 0 ? -0.25f : s == 1 ? 0.00f : s == 2 ? 0.10f : s == 3 ? 0.20f
                    : s == 4 ? 0.40f : s == 5 ? 0.80f : s == 6 ? 1.20f : 1.60f;
    for(int cyc = 0; cyc < 5; ++cyc)
    {
        if(__ballot(active) == 0)
            break;
        const float sx = s == 0 ? -nodeShift : (s == 1 ? nodeShift : 0.0f);
        const float sy = s == 2 ? -nodeShift : (s == 3 ? nodeShift : 0.0f);
        const float sz = s == 4 ? -nodeShift : (s == 5 ? nodeShift : 0.0f);
        const float q = mfmaPatchQuality<EK>(mySp, nRows, nE, x + mk<float>(sx, sy, sz), constSum);
        if(active) evals += 6;
        const float q0 = __shfl(q, base + 0, 64), q1 = __shfl(q, base + 1, 64), q2 = __shfl(q, base + 2, 64);
        const float q3 = __shfl(q, base + 3, 64), q4 = __shfl(q, base + 4, 64), q5 = __shfl(q, base + 5, 64);
        const v3<float> gq = mk<float>(q1 - q0, q3 - q2, q5 - q4);
        const float gn = gmq::tsqrt(gmq::dot(gq, gq));
        if(gn == 0.0f)
            active = false;
        const v3<float> shift = gq * (nodeShift / gn);
        const v3<float> prop = x + shift * off;
        const float qp = mfmaPatchQuality<EK>(mySp, nRows, nE, prop, constSum);
        const int best = groupArgmaxFast<8>(qp);
        const v3<float> nx = shflv<float>(prop, (base & 63) + best);
        const float bestOff = __shfl(off, (base & 63) + best, 64);
        if(active)
        {
            evals += 8;
            x = nx;
            nodeShift *= gmq::tabs(bestOff);
            if(nodeShift < originalNodeShift / 10.0f)
                active = false;
        }
    }
    if(!present)
        return;
    const v3<float> np = testedMove<float, EK>(m, mySp, nE, s, 8, (threadIdx.x & 63) & ~7, o, x);
    if(s == 0)
    {
        st(pos, v, np);
        m.evalCounts[v] += evals;
    }
}

// The nine terms of a tetrahedron corner row that the candidate loop reads
// (36 bytes in LDS instead of a whole 48-byte Corner).
struct RowCoef
{
    float d0, gx, gy, gz, c0, hx, hy, hz, s;
};

// mfmaRowSum over RowCoef rows.
__device__ __forceinline__ float coefRowSum(const RowCoef* rows, int nRows, v3<float> x, float xx)
{
    const int sub = threadIdx.x & 3;
    float sum = 0.0f;
    for(int r0 = 0; r0 < nRows; r0 += 4)
    {
        const RowCoef& k = rows[r0 + sub];
        mfv4 d = {0.0f, 0.0f, 0.0f, 0.0f}, f = {0.0f, 0.0f, 0.0f, 0.0f};
        d = mfma4x4(k.d0, 1.0f, d);
        f = mfma4x4(k.c0, 1.0f, f);
        d = mfma4x4(k.gx, x.x, d);
        f = mfma4x4(k.hx, x.x, f);
        d = mfma4x4(k.gy, x.y, d);
        f = mfma4x4(k.hy, x.y, f);
        d = mfma4x4(k.gz, x.z, d);
        f = mfma4x4(k.hz, x.z, f);
        f = mfma4x4(k.s, xx, f);
#pragma unroll
        for(int r = 0; r < 4; ++r) sum += f[r] * gma::powm23(d[r]);
    }
    return sum;
}

__device__ __forceinline__ gma::Corner<float> coefCorner(const RowCoef& r)
{
    gma::Corner<float> c{};
    c.d0 = r.d0; c.gx = r.gx; c.gy = r.gy; c.gz = r.gz;
    c.c0 = r.c0; c.hx = r.hx; c.hy = r.hy; c.hz = r.hz; c.s = r.s;
    return c;
}

// Gradient descent on tetrahedral patches of at most 8 RPL8 elements: the
// kernel of gdMfmaKernel with 36-byte LDS rows.  The rows' near-degenerate
// thresholds and element references, which only the final inversion test
// reads, stay in the registers of the lane that staged the row (lane s of a
// vertex stages rows s, s + 8, ... and tests the same rows), so more
// wavefronts fit in LDS.
template<int RPL8>
__global__ void __launch_bounds__(512) gdMfmaTetKernel(const DevMesh* __restrict__ meshes,
                                                       const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                       int cap)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    const int g = lane >> 3, s = lane & 7, base = lane & ~7;
    RowCoef* rows = (RowCoef*)smem + (size_t)g * cap;
    const uint32_t i = xcdBlock() * (blockDim.x >> 3) + g;
    const uint4 span = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[span.x];
    const uint32_t item = i < count && !runStopped(m) ? spanVertex(m, span, i) : NO_ITEM;
    const bool present = item != NO_ITEM;
    const uint32_t v = present ? item : 0;
    float4* pos = m.pos32;
    const uint32_t eb = m.eoff[v];
    const int nE = present ? (int)(m.eoff[v + 1] - eb) : 0;
    const v3<float> o = present ? ld<float>(pos, v) : mk<float>(0.0f, 0.0f, 0.0f);
    float thrR[RPL8];
    uint32_t refR[RPL8];
#pragma unroll
    for(int k = 0; k < RPL8; ++k)
    {
        const int e = s + 8 * k;
        thrR[k] = 0.0f;
        refR[k] = 0;
        if(e < nE)
        {
            const uint32_t ref = m.eref[eb + e];
            const uint4 t = m.tets[ref >> 4];
            v3<float> u[4] = {ld<float>(pos, t.x) - o, ld<float>(pos, t.y) - o, ld<float>(pos, t.z) - o,
                              ld<float>(pos, t.w) - o};
            const gma::Corner<float> c = gma::tetCorner<float>(u, (int)(ref & 7), ref);
            rows[e] = RowCoef{c.d0, c.gx, c.gy, c.gz, c.c0, c.hx, c.hy, c.hz, c.s};
            thrR[k] = c.thr;
            refR[k] = ref;
        }
    }
    // Neutral rows (d = 1, f2 = 0) up to the wavefront's largest patch,
    // rounded to 4 rows; the tile holds them.
    int nEmax = nE;
#pragma unroll
    for(int off = 8; off < 64; off <<= 1)
        nEmax = max(nEmax, __shfl_xor(nEmax, off, 64));
    const int nPad = (nEmax + 3) & ~3;
    for(int e = nE + s; e < nPad; e += 8)
        rows[e] = RowCoef{1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    __syncthreads();
    // Patch quality at offset x: 3 n / sum, or the exact rule when a corner
    // is not positive.
    auto quality = [&](v3<float> x) -> float {
        const float xx = gmq::dot(x, x);
        const float sum = coefRowSum(rows, nPad, x, xx);
        if(!gma::fastBad<float>(sum))
            return 3.0f * (float)nE * __builtin_amdgcn_rcpf(sum);
        gma::Acc<float> a;
        for(int e = 0; e < nE; ++e) gma::addTet<float>(a, coefCorner(rows[e]), x, xx);
        return a.finish();
    };

    v3<float> x = mk<float>(0.0f, 0.0f, 0.0f);
    float nodeShift = localSizeGroup<8>(m, pos, v, o, s) * (1.0f / 75.0f);
    const float originalNodeShift = nodeShift;
    int evals = 0;
    bool active = present;
    const float off = s == 0 ? -0.25f : s == 1 ? 0.00f : s == 2 ? 0.10f : s == 3 ? 0.20f
                    : s == 4 ? 0.40f : s == 5 ? 0.80f : s == 6 ? 1.20f : 1.60f;
    for(int cyc = 0; cyc < 5; ++cyc)
    {
        if(__ballot(active) == 0)
            break;
        const float sx = s == 0 ? -nodeShift : (s == 1 ? nodeShift : 0.0f);
        const float sy = s == 2 ? -nodeShift : (s == 3 ? nodeShift : 0.0f);
        const float sz = s == 4 ? -nodeShift : (s == 5 ? nodeShift : 0.0f);
        const float q = quality(x + mk<float>(sx, sy, sz));
        if(active) evals += 6;
        const float q0 = __shfl(q, base + 0, 64), q1 = __shfl(q, base + 1, 64), q2 = __shfl(q, base + 2, 64);
        const float q3 = __shfl(q, base + 3, 64), q4 = __shfl(q, base + 4, 64), q5 = __shfl(q, base + 5, 64);
        const v3<float> gq = mk<float>(q1 - q0, q3 - q2, q5 - q4);
        const float gn = gmq::tsqrt(gmq::dot(gq, gq));
        if(gn == 0.0f)
            active = false;
        const v3<float> shift = gq * (nodeShift / gn);
        const v3<float> prop = x + shift * off;
        const float qp = quality(prop);
        const int best = groupArgmaxFast<8>(qp);
        const v3<float> nx = shflv<float>(prop, (base & 63) + best);
        const float bestOff = __shfl(off, (base & 63) + best, 64);
        if(active)
        {
            evals += 8;
            x = nx;
            nodeShift *= gmq::tabs(bestOff);
            if(nodeShift < originalNodeShift / 10.0f)
                active = false;
        }
    }
    if(!present)
        return;
    // FP64 inversion test of the move over this lane's rows (testedMove).
    const v3<float> pf = o + x;
    const DetFix<0> fix{&m, m.pos32};
    int flags = 0;
#pragma unroll
    for(int k = 0; k < RPL8; ++k)
    {
        const int e = s + 8 * k;
        if(e < nE)
        {
            gma::Corner<float> c = coefCorner(rows[e]);
            c.thr = thrR[k];
            c.ref = refR[k];
            flags |= cornerInversion<0>(c, x, pf, fix);
        }
    }
    const int first = (threadIdx.x & 63) & ~7;
    const bool reject = ((__ballot((flags & INV_FLIP) != 0) >> first) & 0xFFull) != 0 &&
                        ((__ballot((flags & INV_NONPOS32) != 0) >> first) & 0xFFull) == 0;
    if(s == 0)
    {
        st(pos, v, reject ? o : pf);
        m.evalCounts[v] += evals;
    }
}

// Nelder-Mead on affine terms: one vertex per group of G lanes, lane j holding
// element j's staged corner terms in registers; a patch evaluation is one
// affine element evaluation per lane and a butterfly over the group.
template<typename T, int EK, int G>
struct AffineWavePatch
{
    int lane;
    int nElem;
    int count;
    v3<T> o;
    gma::Corner<T> cn;   // EK == 0: in registers
    gma::Hex<T>* hx;     // EK == 1: this lane's slot in LDS

    __device__ T operator()(v3<T> p)
    {
        ++count;
        const v3<T> x = p - o;
        const T xx = gmq::dot(x, x);
        T sum = T(0);
        if(lane < nElem)
            sum = EK == 0 ? gma::fastTet<T>(cn, x, xx) : gma::fastHex<T>(*hx, x, xx);
        if constexpr(sizeof(T) == 4)
            sum = groupSum<G>(sum);
        else
            for(int off = G / 2; off > 0; off >>= 1)
                sum += __shfl_xor(sum, off, 64);
        if(!gma::fastBad<T>(sum))
        {
            if constexpr(sizeof(T) == 4)
                return 3.0f * (float)nElem * __builtin_amdgcn_rcpf(sum);
            else
                return T(3) * (T)nElem / sum;
        }
        // A non-positive element: the smallest non-positive quality.
        gma::Acc<T> a;
        if(lane < nElem)
        {
            if(EK == 0) gma::addTet<T>(a, cn, x, xx);
            else gma::addHex<T>(a, *hx, x, xx);
        }
        T mn = (a.bad && a.mn == a.mn) ? a.mn : (T)__builtin_huge_val();
#pragma unroll
        for(int off = G / 2; off > 0; off >>= 1)
        {
            const T om = __shfl_xor(mn, off, 64);
            mn = om < mn ? om : mn;
        }
        return mn;
    }
};

template<typename T, int EK, int G>
__global__ void __launch_bounds__(64) nmAffineWaveKernel(const DevMesh* __restrict__ meshes,
                                                         const uint4* __restrict__ segs, uint32_t nsegs,
                                                         uint32_t count)
{
    typedef typename PosOf<T>::type P;
    // Hexahedra are staged in LDS, one slot per lane (their moving corners
    // land in vertex-dependent slots, which registers cannot index).
    __shared__ __align__(16) unsigned char hexStage[EK == 1 ? 64 * sizeof(gma::Hex<T>) : 16];
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const uint32_t i = blockIdx.x * (64 / G) + grp;
    if(i >= count)
        return;
    const uint4 sp = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[sp.x];
    const uint32_t v = spanVertex(m, sp, i);
    if(v == NO_ITEM || runStopped(m))
        return;
    P* pos = PosOf<T>::get(m);
    const uint32_t eb = m.eoff[v], ee = m.eoff[v + 1];
    AffineWavePatch<T, EK, G> f;
    f.hx = (gma::Hex<T>*)hexStage + (EK == 1 ? threadIdx.x : 0);
    f.lane = lane;
    f.nElem = (int)(ee - eb);
    f.count = 0;
    f.o = ld<T>(pos, v);
    if(lane < f.nElem)
        stageAffine<T, EK>(m, pos, m.eref[eb + lane], f.o,
                           EK == 0 ? (unsigned char*)&f.cn : (unsigned char*)f.hx);
    const T ls = localSizeGroup<G>(m, pos, v, f.o, lane);
    const v3<T> np = gms::nelderMead<T>(f.o, ls, f);
    // FP64 inversion test of an FP32 move, one element per lane of the group.
    v3<T> out = np;
    if constexpr(sizeof(T) == 4)
    {
        int flags = 0;
        if(lane < f.nElem)
        {
            const DetFix<EK> fix{&m, m.pos32};
            flags = inversionTest<EK>(EK == 0 ? (const unsigned char*)&f.cn : (const unsigned char*)f.hx, 0,
                                      np - f.o, np, fix);
        }
        const int sh = (threadIdx.x & 63) & ~(G - 1);
        const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1);
        const bool reject = ((__ballot((flags & INV_FLIP) != 0) >> sh) & gmask) != 0 &&
                            ((__ballot((flags & INV_NONPOS32) != 0) >> sh) & gmask) == 0;
        out = reject ? f.o : np;
    }
    if(lane == 0)
    {
        st(pos, v, out);
        m.evalCounts[v] += f.count;
    }
}

// Nelder-Mead, FP32 affine terms of tetrahedral patches, four lanes per
// vertex (16 vertices per wavefront): lane k of a vertex holds rows
// k, k + 4, ... of its patch in registers (neutral rows, d = 1 and f2 = 0,
// past the patch), an evaluation sums its rows and adds the four partial
// sums (DPP inside the quad), and the simplex logic of gms::nelderMead runs in
// all four lanes.  The vertices of a wavefront follow their own paths: the
// loop below takes one simplex evaluation and one cycle per trip for every
// vertex still running, and the second evaluation of a cycle only for the
// vertices whose reflection calls for it, so every vertex makes the
// reference's evaluations in the reference's order.
template<int RPL>
__global__ void __launch_bounds__(256) nmQuadKernel(const DevMesh* __restrict__ meshes,
                                                    const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count)
{
    const int sub = threadIdx.x & 3;
    const uint32_t i = xcdBlock() * (blockDim.x >> 2) + (threadIdx.x >> 2);
    if(i >= count)
        return;
    const uint4 sp = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[sp.x];
    const uint32_t v = spanVertex(m, sp, i);
    if(v == NO_ITEM || runStopped(m))
        return;
    float4* pos = m.pos32;
    const uint32_t eb = m.eoff[v];
    const int nE = (int)(m.eoff[v + 1] - eb);
    const v3<float> o = ld<float>(pos, v);
    gma::Corner<float> rows[RPL];
#pragma unroll
    for(int k = 0; k < RPL; ++k)
    {
        const int e = sub + 4 * k;
        if(e < nE)
        {
            const uint32_t ref = m.eref[eb + e];
            const uint4 t = m.tets[ref >> 4];
            v3<float> u[4] = {ld<float>(pos, t.x) - o, ld<float>(pos, t.y) - o, ld<float>(pos, t.z) - o,
                              ld<float>(pos, t.w) - o};
            rows[k] = gma::tetCorner<float>(u, (int)(ref & 7), ref);
        }
        else
        {
            rows[k] = gma::Corner<float>{};
            rows[k].d0 = 1.0f;
        }
    }
    int evals = 0;
    auto f = [&](v3<float> p) -> float {
        ++evals;
        const v3<float> x = p - o;
        const float xx = gmq::dot(x, x);
        float sum = 0.0f;
#pragma unroll
        for(int k = 0; k < RPL; ++k) sum += gma::fastTet<float>(rows[k], x, xx);
        sum = groupSum<4>(sum);
        if(!gma::fastBad<float>(sum))
            return 3.0f * (float)nE * __builtin_amdgcn_rcpf(sum);
        // A non-positive element: the smallest non-positive quality.
        gma::Acc<float> a;
#pragma unroll
        for(int k = 0; k < RPL; ++k)
            if(sub + 4 * k < nE) gma::addTet<float>(a, rows[k], x, xx);
        return groupMin<4>((a.bad && a.mn == a.mn) ? a.mn : __builtin_huge_valf());
    };

    // gms::nelderMead (NelderMeadSmoother.cpp:100-261), one step per trip.
    const float nodeShift = localSizeGroup<4>(m, pos, v, o, sub) * (1.0f / 24.0f);
    gms::v4<float> vo;
    vo.p = o;
    vo.w = f(o);
    gms::v4<float> s[4];
    s[0].p = o + mk<float>(nodeShift, 0.0f, 0.0f); s[0].w = 0.0f;
    s[1].p = o + mk<float>(0.0f, nodeShift, 0.0f); s[1].w = 0.0f;
    s[2].p = o + mk<float>(0.0f, 0.0f, nodeShift); s[2].w = 0.0f;
    s[3] = vo;
    int cycle = 0;
    bool reset = false;
    int phase = 0;   // 0: evaluate the simplex, 1: cycle, 2: done
    while(__ballot(phase != 2) != 0)
    {
        if(phase == 0)
        {
            for(int p = 0; p < 3; ++p) s[p].w = f(s[p].p);
            if(s[0].w > s[1].w) gms::swap4(s[0], s[1]);
            if(s[1].w > s[2].w) gms::swap4(s[1], s[2]);
            if(s[2].w > s[3].w) gms::swap4(s[2], s[3]);
            if(s[0].w > s[1].w) gms::swap4(s[0], s[1]);
            if(s[1].w > s[2].w) gms::swap4(s[1], s[2]);
            if(s[0].w > s[1].w) gms::swap4(s[0], s[1]);
            phase = 1;
        }
        if(phase == 1)
        {
            const v3<float> c = (float)(1 / 3.0) * (s[1].p + s[2].p + s[3].p);
            v3<float> x = c + 1.0f * (c - s[0].p);
            const float fr = f(x);
            float fv = fr;
            const v3<float> xr = x;
            const bool expand = s[3].w < fr;
            const bool contract = !expand && s[1].w >= fr;
            if(expand)
                x = c + 2.0f * (x - c);
            else if(contract)
                x = fr > s[0].w ? c + 0.5f * (xr - c) : c + 0.5f * (s[0].p - c);
            if(expand || contract)
            {
                const float f2 = f(x);
                fv = f2;
                if(expand && f2 <= fr)
                {
                    x = xr;
                    fv = fr;
                }
            }
            gms::v4<float> vtx;
            vtx.p = x;
            vtx.w = fv;
            if(vtx.w > s[3].w) gms::swap4(s[3], vtx);
            if(vtx.w > s[2].w) gms::swap4(s[2], vtx);
            if(vtx.w > s[1].w) gms::swap4(s[1], vtx);
            if(vtx.w > s[0].w) gms::swap4(s[0], vtx);
            if((s[3].w - s[1].w) < 0.000100f)
                phase = 2;
            else if(++cycle >= 12)
            {
                if(reset)
                    phase = 2;
                else
                {
                    s[0].p = vo.p - mk<float>(nodeShift, 0.0f, 0.0f); s[0].w = vo.w - 0.0f;
                    s[1].p = vo.p - mk<float>(0.0f, nodeShift, 0.0f); s[1].w = vo.w - 0.0f;
                    s[2].p = vo.p - mk<float>(0.0f, 0.0f, nodeShift); s[2].w = vo.w - 0.0f;
                    s[3] = vo;
                    reset = true;
                    cycle = 0;
                    phase = 0;
                }
            }
        }
    }
    const v3<float> np = s[3].p;
    // FP64 inversion test of the move, the quad's rows.
    const DetFix<0> fix{&m, m.pos32};
    int flags = 0;
#pragma unroll
    for(int k = 0; k < RPL; ++k)
        if(sub + 4 * k < nE) flags |= cornerInversion<0>(rows[k], np - o, np, fix);
    const int sh = (threadIdx.x & 63) & ~3;
    const bool reject = ((__ballot((flags & INV_FLIP) != 0) >> sh) & 15ull) != 0 &&
                        ((__ballot((flags & INV_NONPOS32) != 0) >> sh) & 15ull) == 0;
    if(sub == 0)
    {
        st(pos, v, reject ? o : np);
        m.evalCounts[v] += evals;
    }
}

template<typename T, int EK>
static void launchNmAffine(const VertexWork& w, hipStream_t stream)
{
    if constexpr(sizeof(T) == 4 && EK == 0)
    {
        // Four lanes per vertex, 64 vertices per 256-thread workgroup.
        const uint32_t blocks = (w.count + 63) / 64;
        if(w.maxPatch <= 8) { nmQuadKernel<2><<<blocks, 256, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count); return; }
        if(w.maxPatch <= 16) { nmQuadKernel<4><<<blocks, 256, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count); return; }
        if(w.maxPatch <= 24) { nmQuadKernel<6><<<blocks, 256, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count); return; }
        if(w.maxPatch <= 32) { nmQuadKernel<8><<<blocks, 256, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count); return; }
    }
    if(w.maxPatch <= 8)
        nmAffineWaveKernel<T, EK, 8><<<(w.count + 7) / 8, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count);
    else if(w.maxPatch <= 16)
        nmAffineWaveKernel<T, EK, 16><<<(w.count + 3) / 4, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count);
    else if(w.maxPatch <= 32)
        nmAffineWaveKernel<T, EK, 32><<<(w.count + 1) / 2, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count);
    else
        nmAffineWaveKernel<T, EK, 64><<<w.count, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count);
}

// ---------------------------------------------------------------------------
// GETMe element step (Smoothers/ElementWise/GetmeSmoother.cpp:55-127, 245-374):
// every element's transformed positions and weight against the pass-start
// positions; out[n_vertices_per_element * id + k] = (vpp_k, weight).  The
// element's quality goes to getmeQual (tets, then hexes), which the vertex
// steps of the pass keep current.
// ---------------------------------------------------------------------------
template<typename T, int EV, int SA, bool FAST = false>
__global__ void __launch_bounds__(ELEM_BLOCK) getmeTetKernel(const DevMesh* __restrict__ meshes,
                                                             const uint4* __restrict__ spans, uint32_t nspans,
                                                             gmq::Crew c)
{
    // FAST (FP64, mean ratio, uniform sampler): the mixed-precision passes'
    // routines -- seeded cube root and tet quality, as their vertex step.
    constexpr bool FASTE = FAST && sizeof(T) == 8 && EV == 0 && SA == 0;
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    const uint32_t e = bl.y * ELEM_BLOCK + threadIdx.x;
    if(e >= m.ntets || runStopped(m))
        return;
    typedef typename PosOf<T>::type P;
    const P* pos = PosOf<T>::get(m);
    P* out = (P*)m.getmePull;
    const T lambda = T(0.78);
    uint4 t = m.tets[e];
    uint32_t vi[4] = {t.x, t.y, t.z, t.w};
    v3<T> vp[4];
#pragma unroll
    for(int k = 0; k < 4; ++k) vp[k] = ld<T>(pos, vi[k]);
    v3<T> n[4] = {gmq::cross(vp[3] - vp[1], vp[2] - vp[1]), gmq::cross(vp[3] - vp[2], vp[0] - vp[2]),
                  gmq::cross(vp[1] - vp[3], vp[0] - vp[3]), gmq::cross(vp[1] - vp[0], vp[2] - vp[0])};
    v3<T> vpp[4];
#pragma unroll
    for(int k = 0; k < 4; ++k) vpp[k] = vp[k] + n[k] * gmq::tdiv(lambda, gmq::tsqrt(gmq::tsqrt(gmq::dot(n[k], n[k]))));
    T volume = gmq::tetVolume<T, SA>(c, vp);
    T volumePrime = gmq::tetVolume<T, SA>(c, vpp);
    T volumeVar;
    if constexpr(FASTE)
        volumeVar = fastCbrt(gmq::tabs(gmq::tdiv(volume, volumePrime)));
    else
        volumeVar = gmq::tcbrt(gmq::tabs(gmq::tdiv(volume, volumePrime)));
    v3<T> center = T(1.0 / 4.0) * (vp[0] + vp[1] + vp[2] + vp[3]);
#pragma unroll
    for(int k = 0; k < 4; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
#pragma unroll
    for(int k = 0; k < 4; ++k)
    {
        if(m.fixed[vi[k]]) vpp[k] = vp[k];
        else if(const double* cs = consOf(m, vi[k])) vpp[k] = DevSnap<T>{cs}(vpp[k]);
    }
    T quality, qualityPrime;
    if constexpr(FASTE)
    {
        double inv;
        quality = fastTetEval(c, vp, inv);
        qualityPrime = fastTetEval(c, vpp, inv);
    }
    else
    {
        quality = gmq::tetQuality<T, EV, SA>(c, vp);
        qualityPrime = gmq::tetQuality<T, EV, SA>(c, vpp);
    }
    T weight = gmq::tdiv(qualityPrime, T(1) + quality);
    ((T*)m.getmeQual)[e] = quality;
#pragma unroll
    for(int k = 0; k < 4; ++k)
    {
        P o;
        o.x = vpp[k].x; o.y = vpp[k].y; o.z = vpp[k].z; o.w = weight;
        out[m.cornerSlot[4 * (size_t)e + k]] = o;
    }
}

template<typename T, int EV, int SA>
__global__ void __launch_bounds__(ELEM_BLOCK) getmeHexKernel(const DevMesh* __restrict__ meshes,
                                                             const uint4* __restrict__ spans, uint32_t nspans,
                                                             gmq::Crew c)
{
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    const uint32_t e = bl.y * ELEM_BLOCK + threadIdx.x;
    if(e >= m.nhexs || runStopped(m))
        return;
    typedef typename PosOf<T>::type P;
    const P* pos = PosOf<T>::get(m);
    P* out = (P*)m.getmePull;
    const T lambda = T(0.78);
    uint4 ha = m.hexs[2 * e], hb = m.hexs[2 * e + 1];
    uint32_t vi[8] = {ha.x, ha.y, ha.z, ha.w, hb.x, hb.y, hb.z, hb.w};
    v3<T> vp[8];
#pragma unroll
    for(int k = 0; k < 8; ++k) vp[k] = ld<T>(pos, vi[k]);
    const T q4 = T(4), t3 = T(3);
    v3<T> aux[6] = {(vp[0] + vp[1] + vp[2] + vp[3]) / q4, (vp[0] + vp[4] + vp[5] + vp[1]) / q4,
                    (vp[1] + vp[5] + vp[6] + vp[2]) / q4, (vp[2] + vp[6] + vp[7] + vp[3]) / q4,
                    (vp[0] + vp[3] + vp[7] + vp[4]) / q4, (vp[4] + vp[7] + vp[6] + vp[5]) / q4};
    v3<T> n[8] = {gmq::cross(aux[1] - aux[0], aux[4] - aux[0]), gmq::cross(aux[2] - aux[0], aux[1] - aux[0]),
                  gmq::cross(aux[3] - aux[0], aux[2] - aux[0]), gmq::cross(aux[4] - aux[0], aux[3] - aux[0]),
                  gmq::cross(aux[4] - aux[5], aux[1] - aux[5]), gmq::cross(aux[1] - aux[5], aux[2] - aux[5]),
                  gmq::cross(aux[2] - aux[5], aux[3] - aux[5]), gmq::cross(aux[3] - aux[5], aux[4] - aux[5])};
    v3<T> bases[8] = {(aux[0] + aux[1] + aux[4]) / t3, (aux[0] + aux[2] + aux[1]) / t3,
                      (aux[0] + aux[3] + aux[2]) / t3, (aux[0] + aux[4] + aux[3]) / t3,
                      (aux[5] + aux[4] + aux[1]) / t3, (aux[5] + aux[1] + aux[2]) / t3,
                      (aux[5] + aux[2] + aux[3]) / t3, (aux[5] + aux[3] + aux[4]) / t3};
    v3<T> vpp[8];
#pragma unroll
    for(int k = 0; k < 8; ++k) vpp[k] = bases[k] + n[k] * gmq::tdiv(lambda, gmq::tsqrt(gmq::tsqrt(gmq::dot(n[k], n[k]))));
    T volume = gmq::hexVolume<T, SA>(c, vp);
    T volumePrime = gmq::hexVolume<T, SA>(c, vpp);
    T volumeVar = gmq::tcbrt(gmq::tabs(gmq::tdiv(volume, volumePrime)));
    v3<T> center = T(1.0 / 8.0) * (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5] + vp[6] + vp[7]);
#pragma unroll
    for(int k = 0; k < 8; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
#pragma unroll
    for(int k = 0; k < 8; ++k)
    {
        if(m.fixed[vi[k]]) vpp[k] = vp[k];
        else if(const double* cs = consOf(m, vi[k])) vpp[k] = DevSnap<T>{cs}(vpp[k]);
    }
    T quality = gmq::hexQuality<T, EV, SA>(c, vp);
    T qualityPrime = gmq::hexQuality<T, EV, SA>(c, vpp);
    T weight = gmq::tdiv(qualityPrime, T(1) + quality);
    ((T*)m.getmeQual)[m.ntets + e] = quality;
#pragma unroll
    for(int k = 0; k < 8; ++k)
    {
        P o;
        o.x = vpp[k].x; o.y = vpp[k].y; o.z = vpp[k].z; o.w = weight;
        out[m.cornerSlot[4 * (size_t)m.ntets + 8 * (size_t)e + k]] = o;
    }
}

// GetmeSmoother::smoothPris (GetmeSmoother.cpp:131-243): prism e is record
// nhexs + e.  Base points sit between the triangle centres and the side-face
// centres at t = (4/5)(1 - (4/39)^(1/4) lambda), which maps the regular prism
// onto itself.
template<typename T, int EV, int SA>
__global__ void __launch_bounds__(ELEM_BLOCK) getmePriKernel(const DevMesh* __restrict__ meshes,
                                                             const uint4* __restrict__ spans, uint32_t nspans,
                                                             gmq::Crew c)
{
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    const uint32_t e = bl.y * ELEM_BLOCK + threadIdx.x;
    if(e >= m.npris || runStopped(m))
        return;
    typedef typename PosOf<T>::type P;
    const P* pos = PosOf<T>::get(m);
    P* out = (P*)m.getmePull;
    const T lambda = T(0.78);
    const size_t r = (size_t)m.nhexs + e;
    uint4 ha = m.hexs[2 * r], hb = m.hexs[2 * r + 1];
    uint32_t vi[6] = {ha.x, ha.y, ha.z, ha.w, hb.x, hb.y};
    v3<T> vp[6];
#pragma unroll
    for(int k = 0; k < 6; ++k) vp[k] = ld<T>(pos, vi[k]);
    const T q4 = T(4), t3 = T(3), two = T(2);
    v3<T> aux[5] = {(vp[0] + vp[1] + vp[2]) / t3, (vp[0] + vp[1] + vp[4] + vp[3]) / q4,
                    (vp[1] + vp[2] + vp[5] + vp[4]) / q4, (vp[2] + vp[0] + vp[3] + vp[5]) / q4,
                    (vp[3] + vp[4] + vp[5]) / t3};
    v3<T> n[6] = {gmq::cross(aux[1] - aux[0], aux[3] - aux[0]), gmq::cross(aux[2] - aux[0], aux[1] - aux[0]),
                  gmq::cross(aux[3] - aux[0], aux[2] - aux[0]), gmq::cross(aux[3] - aux[4], aux[1] - aux[4]),
                  gmq::cross(aux[1] - aux[4], aux[2] - aux[4]), gmq::cross(aux[2] - aux[4], aux[3] - aux[4])};
    const T t = (T)((4.0 / 5.0) * (1.0 - gm_pow(4.0 / 39.0, 0.25) * 0.78));
    const T it = T(1) - t;
    v3<T> bases[6] = {it * aux[0] + t * (aux[3] + aux[1]) / two, it * aux[0] + t * (aux[1] + aux[2]) / two,
                      it * aux[0] + t * (aux[2] + aux[3]) / two, it * aux[4] + t * (aux[3] + aux[1]) / two,
                      it * aux[4] + t * (aux[1] + aux[2]) / two, it * aux[4] + t * (aux[2] + aux[3]) / two};
    v3<T> vpp[6];
#pragma unroll
    for(int k = 0; k < 6; ++k) vpp[k] = bases[k] + n[k] * gmq::tdiv(lambda, gmq::tsqrt(gmq::tsqrt(gmq::dot(n[k], n[k]))));
    T volume = gmq::priVolume<T, SA>(c, vp);
    T volumePrime = gmq::priVolume<T, SA>(c, vpp);
    T volumeVar = gmq::tcbrt(gmq::tabs(gmq::tdiv(volume, volumePrime)));
    v3<T> center = T(1.0 / 6.0) * (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5]);
#pragma unroll
    for(int k = 0; k < 6; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
#pragma unroll
    for(int k = 0; k < 6; ++k)
    {
        if(m.fixed[vi[k]]) vpp[k] = vp[k];
        else if(const double* cs = consOf(m, vi[k])) vpp[k] = DevSnap<T>{cs}(vpp[k]);
    }
    T quality = gmq::priQuality<T, EV, SA>(c, vp);
    T qualityPrime = gmq::priQuality<T, EV, SA>(c, vpp);
    T weight = gmq::tdiv(qualityPrime, T(1) + quality);
    ((T*)m.getmeQual)[m.ntets + r] = quality;
#pragma unroll
    for(int k = 0; k < 6; ++k)
    {
        P o;
        o.x = vpp[k].x; o.y = vpp[k].y; o.z = vpp[k].z; o.w = weight;
        out[m.cornerSlot[4 * (size_t)m.ntets + 8 * r + k]] = o;
    }
}

// GETMe vertex step (AbstractElementWiseSmoother::updateVertexPositions,
// AbstractElementWiseSmoother.cpp:406-444; accumulation VertexAccum.cpp:28-49)
// for one colour group: pull the contributions in patch order (the push order
// of the element loop), move to their weighted mean, keep the move unless the
// patch quality drops.
template<typename T, int EV, int SA>
__global__ void __launch_bounds__(512) getmeVertexKernel(const DevMesh* __restrict__ meshes,
                                                         const uint4* __restrict__ segs, uint32_t nsegs, uint32_t count,
                                                         gmq::Crew c)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count)
        return;
    typedef typename PosOf<T>::type P;
    const uint2 it = workItem(meshes, segs, nsegs, i);
    if(it.y == NO_ITEM || runStopped(meshes[it.x]))
        return;
    const DevMesh& m = meshes[it.x];
    P* pos = PosOf<T>::get(m);
    const P* pull = (const P*)m.getmePull;
    const uint32_t v = it.y;
    v3<T> acc = mk<T>(T(0), T(0), T(0));
    T wacc = T(0);
    uint32_t b = m.eoff[v], e = m.eoff[v + 1];
    for(uint32_t r = b; r < e; ++r)
    {
        P o = pull[r];
        T w = (T)o.w;
        acc = acc + mk<T>((T)o.x, (T)o.y, (T)o.z) * w;
        wacc += w;
    }
    if(wacc != T(0))
    {
        v3<T> p0 = ld<T>(pos, v);
        v3<T> p1 = DevSnap<T>{consOf(m, v)}(acc / wacc);
        PlainPatch<T, EV, SA> f{m, c, pos, v, 0};
        T q0 = f(p0);
        T q1 = f(p1);
        if(!(q1 < q0))
            st(pos, v, p1);
        m.evalCounts[v] += f.count;
    }
}

// GETMe vertex step, one vertex per group of G lanes (G >= patch size): lane j
// pulls element j's contribution and holds its vertices; in FP64 the
// weighted sum is formed in patch order through LDS (bit-identical to
// getmeVertexKernel), in FP32 by a DPP tree.
// The patch quality before the move is reduced from getmeQual: every element
// quality there was computed from the element's current vertex positions
// (by the element step, or by the accepted move that last changed one of
// them), so it has the bits a fresh evaluation would give.  The quality after
// the move is evaluated lane-parallel as in smoothWaveKernel and, when the
// move is kept, written back.  GETME_WAVES wavefronts per workgroup: more
// resident wavefronts to cover the dependent gathers.
constexpr int GETME_WAVES = 4;

template<typename T, int EV, int SA, int G, int EKIND, bool FASTQ>
__global__ void __launch_bounds__(64 * GETME_WAVES) getmeWaveKernel(const DevMesh* __restrict__ meshes,
                                                                    const uint4* __restrict__ segs, uint32_t nsegs,
                                                                    uint32_t count, gmq::Crew c)
{
    typedef typename PosOf<T>::type P;
    __shared__ T red[2 * 64 * GETME_WAVES];
    __shared__ T pull[sizeof(T) == 8 ? 4 * 64 * GETME_WAVES : 1];
    const int lane = threadIdx.x % G;
    const int grp = threadIdx.x / G;
    const uint32_t i = xcdBlock() * (64 * GETME_WAVES / G) + grp;
    if(i >= count)
        return;
    const uint4 sp = waveSpan(segs, nsegs, i);
    const DevMesh& m = meshes[sp.x];
    const uint32_t v = spanVertex(m, sp, i);
    if(v == NO_ITEM || runStopped(m))
        return;
    P* pos = PosOf<T>::get(m);
    const P* pullBuf = (const P*)m.getmePull;
    WavePatch<T, EV, SA, G, EKIND, FASTQ> f;
    f.c = &c;
    f.red = red + 2 * G * grp;
    f.lane = lane;
    f.count = 0;
    const uint32_t eb = m.eoff[v], ee = m.eoff[v + 1];
    f.nElem = (int)(ee - eb);
    f.isHex = 0;
    f.isPri = 0;
    f.k = 0;
    T* pl = pull + 4 * G * grp;
    T* qual = (T*)m.getmeQual;
    uint32_t qslot = 0;
    T qOld = T(1);
    v3<T> pw = mk<T>(T(0), T(0), T(0));
    T w = T(0);
    if(lane < f.nElem)
    {
        const uint32_t ref = m.eref[eb + lane];
        const uint32_t id = ref >> 4;
        f.k = (int)(ref & 7);
        f.isHex = (ref >> 3) & 1;
        f.isPri = f.isHex && id >= m.nhexs;
        qslot = f.isHex ? m.ntets + id : id;
        qOld = qual[qslot];
        P o;
        if(EKIND == EK_TET || (EKIND == EK_MIXED && !f.isHex))
        {
            uint4 t = m.tets[id];
            f.vp[0] = ld<T>(pos, t.x); f.vp[1] = ld<T>(pos, t.y);
            f.vp[2] = ld<T>(pos, t.z); f.vp[3] = ld<T>(pos, t.w);
            o = pullBuf[eb + lane];
        }
        else if constexpr(EKIND != EK_TET)
        {
            uint4 a = m.hexs[2 * id], b = m.hexs[2 * id + 1];
            f.vp[0] = ld<T>(pos, a.x); f.vp[1] = ld<T>(pos, a.y);
            f.vp[2] = ld<T>(pos, a.z); f.vp[3] = ld<T>(pos, a.w);
            f.vp[4] = ld<T>(pos, b.x); f.vp[5] = ld<T>(pos, b.y);
            f.vp[6] = ld<T>(pos, b.z); f.vp[7] = ld<T>(pos, b.w);
            o = pullBuf[eb + lane];
        }
        w = (T)o.w;
        pw = mk<T>((T)o.x, (T)o.y, (T)o.z) * w;
    }
    v3<T> acc = mk<T>(T(0), T(0), T(0));
    T wacc = T(0);
    if constexpr(sizeof(T) == 4 || FASTQ)
    {
        // FP32 and the mixed-precision FP64 passes: a tree over the group
        // (lanes past the patch add zeros).
        acc = mk<T>(groupSum<G>(pw.x), groupSum<G>(pw.y), groupSum<G>(pw.z));
        wacc = groupSum<G>(w);
    }
    else
    {
        // FP64: the patch order of the reference's accumulation.
        if(lane < f.nElem)
        {
            pl[4 * lane] = pw.x;
            pl[4 * lane + 1] = pw.y;
            pl[4 * lane + 2] = pw.z;
            pl[4 * lane + 3] = w;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_wave_barrier();
        for(int j = 0; j < f.nElem; ++j)
        {
            acc = acc + mk<T>(pl[4 * j], pl[4 * j + 1], pl[4 * j + 2]);
            wacc += pl[4 * j + 3];
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_wave_barrier();
    }
    if(wacc != T(0))
    {
        const v3<T> p1 = DevSnap<T>{consOf(m, v)}(acc / wacc);
        const T q0 = f.reduce(qOld);
        const T qNew = f.elemQuality(p1);
        const T q1 = f.reduce(qNew);
        const bool keep = !(q1 < q0);
        if(keep && lane < f.nElem)
            qual[qslot] = qNew;
        if(lane == 0)
        {
            if(keep)
                st(pos, v, p1);
            m.evalCounts[v] += 2;
        }
    }
}

// ---------------------------------------------------------------------------
// Mesh quality (AbstractEvaluator::evaluateMeshQuality*, AbstractEvaluator.cpp:384-453;
// QualityHistogram::add, DataStructures/QualityHistogram.cpp:63-92) in FP64.
// Each block owns a contiguous element range; per-lane compensated sums are
// combined in a fixed tree, so the result is deterministic.
// ---------------------------------------------------------------------------
template<int EV, int SA, typename P, bool HEX>
__global__ void __launch_bounds__(EVAL_THREADS) evalKernel(const DevMesh* __restrict__ meshes,
                                                           const uint4* __restrict__ spans, uint32_t nspans,
                                                           gmq::Crew c)
{
    __shared__ double smin[EVAL_THREADS], ssum[EVAL_THREADS], scmp[EVAL_THREADS];
    __shared__ int shist[20];
    __shared__ int sbad;
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    if(runStopped(m))
        return;
    if(threadIdx.x < 20) shist[threadIdx.x] = 0;
    if(threadIdx.x == 0) sbad = 0;
    __syncthreads();
    const P* pos = PosOf<typename std::conditional<std::is_same<P, double4>::value, double, float>::type>::get(m);
    // The tet-only variant never reads past the tetrahedra, whatever it is
    // launched on.
    const uint64_t E = HEX ? (uint64_t)m.ntets + m.nhexs + m.npris : (uint64_t)m.ntets;
    const uint64_t chunk = (E + m.evalBlocks - 1) / m.evalBlocks;
    const uint64_t b0 = (uint64_t)bl.y * chunk;
    const uint64_t b1 = (b0 + chunk < E) ? b0 + chunk : E;
    // FP32 records of mean-ratio metric-free runs take the fast FP64 routines.
    constexpr bool FAST = EV == 0 && SA == 0 && std::is_same<P, float4>::value;
    double mn = 1.0, sum = 0.0, cmp = 0.0;
    int bad = 0;
    for(uint64_t e = b0 + threadIdx.x; e < b1; e += blockDim.x)
    {
        double q, inv = 0.0;
        // Per-element qualities go out in the reference's order: tets, then
        // prisms, then hexahedra (records hold the hexahedra first).
        uint64_t out = e;
        if(!HEX || e < m.ntets)
        {
            uint4 t = m.tets[e];
            v3<double> vp[4] = {ld<double>(pos, t.x), ld<double>(pos, t.y), ld<double>(pos, t.z), ld<double>(pos, t.w)};
            if constexpr(FAST)
                q = fastTetEval(c, vp, inv);
            else
                q = gmq::tetQuality<double, EV, SA>(c, vp);
        }
        else if constexpr(HEX)
        {
            uint64_t h = e - m.ntets;
            uint4 a = m.hexs[2 * h], bq = m.hexs[2 * h + 1];
            v3<double> vp[8] = {ld<double>(pos, a.x), ld<double>(pos, a.y), ld<double>(pos, a.z), ld<double>(pos, a.w),
                                ld<double>(pos, bq.x), ld<double>(pos, bq.y), ld<double>(pos, bq.z), ld<double>(pos, bq.w)};
            if(h >= m.nhexs)
            {
                q = gmq::priQuality<double, EV, SA>(c, vp);
                inv = 1.0 / q;
                out = m.ntets + (h - m.nhexs);
            }
            else
            {
                if constexpr(FAST)
                    q = fastHexEval(c, vp, inv);
                else
                    q = gmq::hexQuality<double, EV, SA>(c, vp);
                out = (uint64_t)m.ntets + m.npris + h;
            }
        }
        if(m.elemQual) m.elemQual[out] = q;
        mn = gmq::tmin(mn, q);
        if(q > 0.0)
        {
            double t, err;
            twoSum(sum, FAST ? inv : 1.0 / q, t, err);
            sum = t;
            cmp += err;
        }
        else
            bad = 1;
        // One LDS increment per distinct bucket of the wavefront (most
        // elements share a bucket; per-lane atomics on one address
        // serialise).
        const int bucket = histBucket(q);
        unsigned long long pending = __ballot(true);
        while(pending)
        {
            const int leader = (int)__builtin_ctzll(pending);
            const int lb = __builtin_amdgcn_readlane(bucket, leader);
            const unsigned long long same = __ballot(bucket == lb);
            if((int)(threadIdx.x & 63) == leader)
                atomicAdd(&shist[lb], (int)__builtin_popcountll(same));
            pending &= ~same;
        }
    }
    smin[threadIdx.x] = mn;
    ssum[threadIdx.x] = sum;
    scmp[threadIdx.x] = cmp;
    if(bad) sbad = 1;
    __syncthreads();
    for(int w = EVAL_THREADS / 2; w > 0; w >>= 1)
    {
        if(threadIdx.x < (unsigned)w)
        {
            double a = smin[threadIdx.x], bb = smin[threadIdx.x + w];
            smin[threadIdx.x] = gmq::tmin(a, bb);
            double t, err;
            twoSum(ssum[threadIdx.x], ssum[threadIdx.x + w], t, err);
            ssum[threadIdx.x] = t;
            scmp[threadIdx.x] = scmp[threadIdx.x] + scmp[threadIdx.x + w] + err;
        }
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        m.evalPartials[3 * bl.y] = smin[0];
        m.evalPartials[3 * bl.y + 1] = ssum[0];
        m.evalPartials[3 * bl.y + 2] = scmp[0];
        m.evalFlags[bl.y] = sbad;
    }
    if(threadIdx.x < 20 && shist[threadIdx.x])
        atomicAdd(&m.evalHist[threadIdx.x], shist[threadIdx.x]);
}

// One block per mesh: reduces the mesh's block partials in a fixed order and
// writes pass record *passSlot, then advances the slot.
static __global__ void __launch_bounds__(EVAL_THREADS) evalFinalKernel(const DevMesh* __restrict__ meshes,
                                                                       const uint32_t* __restrict__ meshList)
{
    __shared__ double smin[EVAL_THREADS], ssum[EVAL_THREADS], scmp[EVAL_THREADS];
    __shared__ int sbad;
    const DevMesh& m = meshes[meshList[blockIdx.x]];
    const int slot = m.passSlot[0];
    // The host sizes the record buffers for the run; a slot past them is
    // never written.
    if(slot < 0 || (uint32_t)slot >= m.resultCap || runStopped(m))
        return;
    if(threadIdx.x == 0) sbad = 0;
    __syncthreads();
    double mn = 1.0, sum = 0.0, cmp = 0.0;
    for(uint32_t b = threadIdx.x; b < m.evalBlocks; b += blockDim.x)
    {
        mn = gmq::tmin(mn, m.evalPartials[3 * b]);
        double t, err;
        twoSum(sum, m.evalPartials[3 * b + 1], t, err);
        sum = t;
        cmp += err + m.evalPartials[3 * b + 2];
        if(m.evalFlags[b]) sbad = 1;
    }
    smin[threadIdx.x] = mn;
    ssum[threadIdx.x] = sum;
    scmp[threadIdx.x] = cmp;
    __syncthreads();
    for(int w = EVAL_THREADS / 2; w > 0; w >>= 1)
    {
        if(threadIdx.x < (unsigned)w)
        {
            smin[threadIdx.x] = gmq::tmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            double t, err;
            twoSum(ssum[threadIdx.x], ssum[threadIdx.x + w], t, err);
            ssum[threadIdx.x] = t;
            scmp[threadIdx.x] = scmp[threadIdx.x] + scmp[threadIdx.x + w] + err;
        }
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        const double mn0 = smin[0];
        const double inv = sbad ? (double)__builtin_huge_val() : ssum[0] + scmp[0];
        const double n = (double)m.ntets + (double)m.nhexs + (double)m.npris;
        m.evalResult[3 * slot] = mn0;
        m.evalResult[3 * slot + 1] = inv;
        m.evalResult[3 * slot + 2] = n;
        if(m.autopilot && slot > 0)
        {
            // Gains against the previous record, as AbstractSmoother does
            // (harmonic mean = n / sum of 1/q).
            const double minGain = mn0 - m.evalResult[3 * (slot - 1)];
            const double geomGain = n / inv - m.evalResult[3 * (slot - 1) + 2] / m.evalResult[3 * (slot - 1) + 1];
            if(!((geomGain > m.apMeanGain) || (minGain > m.apMinGain)))
                m.passSlot[1] = 1;
        }
    }
    if(threadIdx.x < 20)
    {
        m.evalHistOut[20 * slot + threadIdx.x] = m.evalHist[threadIdx.x];
        m.evalHist[threadIdx.x] = 0;
    }
    __syncthreads();
    if(threadIdx.x == 0)
        m.passSlot[0] = slot + 1;
}

// ---------------------------------------------------------------------------
// Launchers of one (precision, evaluator, sampler) combination; each
// combination is instantiated in its own translation unit (parallel build)
// and selected at run time by dispatch.hip.
// ---------------------------------------------------------------------------
template<typename T, int EV, int SA, int SM, int EK>
static hipError_t launchLane(const VertexWork& w, const gmq::Crew& c, LaneShape lane, double moveCoeff,
                             hipStream_t stream)
{
    const int ppg = lane.patchesPerGroup;
    const bool aff = useAffine<T, EV, SA>(w);
    if(SM == SM_SPAWN_SEARCH)
    {
        // LDS tile per block = ppg wavefronts x tileElems elements.
        const int cap = lane.tileElems;
        const size_t lds = (size_t)ppg * laneStageBytes<T, EK>(cap, aff);
        const uint32_t blocks = (w.count + ppg - 1) / ppg;
        if constexpr(AffineCrew<EV, SA>::on)
            if(aff)
            {
                if constexpr(sizeof(T) == 4)
                {
                    // SPAWN_VPW vertices (tiles) per wavefront.
                    const uint32_t vb = (w.count + SPAWN_VPW * ppg - 1) / (SPAWN_VPW * ppg);
                    spawnMfmaKernel<EK, SPAWN_VPW><<<vb, 64 * ppg, SPAWN_VPW * lds, stream>>>(
                        w.meshes, w.segs, w.nsegs, w.count, (float)moveCoeff, cap);
                }
                else
                {
                    const uint32_t vb = (w.count + SPAWN_VPW * ppg - 1) / (SPAWN_VPW * ppg);
                    spawnAffineKernel<EK, SPAWN_VPW><<<vb, 64 * ppg, SPAWN_VPW * lds, stream>>>(
                        w.meshes, w.segs, w.nsegs, w.count, moveCoeff, cap);
                }
                return hipGetLastError();
            }
        spawnLaneKernel<T, EV, SA, EK><<<blocks, 64 * ppg, lds, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c,
                                                                          (T)moveCoeff, cap);
    }
    else
    {
        // 8 vertices per wavefront share the wavefront's tile.
        const int cap = lane.tileElems / 8;
        const size_t lds = (size_t)8 * ppg * laneStageBytes<T, EK>(cap, aff);
        const uint32_t blocks = (w.count + 8 * ppg - 1) / (8 * ppg);
        if constexpr(AffineCrew<EV, SA>::on)
            if(aff)
            {
                if constexpr(sizeof(T) == 4)
                {
                    if constexpr(EK == EK_TET)
                        if(cap <= 32)
                        {
                            // 36-byte rows; the threshold and reference of
                            // lane s's rows s + 8 k, k < RPL8, in registers.
                            const size_t tl = (size_t)8 * ppg * cap * sizeof(RowCoef);
                            if(cap <= 8) gdMfmaTetKernel<1><<<blocks, 64 * ppg, tl, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                            else if(cap <= 16) gdMfmaTetKernel<2><<<blocks, 64 * ppg, tl, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                            else if(cap <= 24) gdMfmaTetKernel<3><<<blocks, 64 * ppg, tl, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                            else gdMfmaTetKernel<4><<<blocks, 64 * ppg, tl, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                            return hipGetLastError();
                        }
                    gdMfmaKernel<EK><<<blocks, 64 * ppg, lds, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                }
                else
                {
                    if constexpr(EK == EK_TET)
                    {
                        // 72-byte rows (the nine terms a candidate reads).
                        const size_t tl = (size_t)8 * ppg * cap * sizeof(RowCoefD);
                        gdAffineTetKernel<<<blocks, 64 * ppg, tl, stream>>>(w.meshes, w.segs, w.nsegs, w.count, cap);
                    }
                    else
                        gdAffineKernel<T, EK><<<blocks, 64 * ppg, lds, stream>>>(w.meshes, w.segs, w.nsegs, w.count,
                                                                                cap);
                }
                return hipGetLastError();
            }
        gdLaneKernel<T, EV, SA, EK><<<blocks, 64 * ppg, lds, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, cap);
    }
    return hipGetLastError();
}

// True when the candidate-parallel kernel of `smoother` can run `w` with `lane`.
template<typename T, int EV, int SA>
inline bool laneFits(const VertexWork& w, int smoother, LaneShape lane)
{
    if(!(smoother == SM_SPAWN_SEARCH || smoother == SM_GRADIENT_DESCENT) || w.kind == EK_MIXED)
        return false;
    const int ppg = lane.patchesPerGroup;
    const int cap = smoother == SM_SPAWN_SEARCH ? lane.tileElems : lane.tileElems / 8;
    if(cap < w.maxPatch || ppg < 1 || ppg > 8)
        return false;
    const bool aff = useAffine<T, EV, SA>(w);
    const size_t per = w.kind == EK_TET ? laneStageBytes<T, EK_TET>(cap, aff) : laneStageBytes<T, EK_HEX>(cap, aff);
    const size_t spawnTiles = (size_t)ppg * (aff ? SPAWN_VPW : 1);
    const size_t lds = (smoother == SM_SPAWN_SEARCH ? spawnTiles : (size_t)8 * ppg) * per;
    return lds <= 64 * 1024;
}

template<typename T, int EV, int SA, int SM>
static hipError_t launchSmoothTyped(const VertexWork& w, const gmq::Crew& c, int shape, LaneShape lane,
                                    double moveCoeff, hipStream_t stream)
{
    if(w.count == 0)
        return hipSuccess;
    // Sliding vertices run on the generic kernels, which snap every trial
    // position; the affine and candidate-parallel kernels take free ones.
    if constexpr(AffineCrew<EV, SA>::on && SM == SM_NELDER_MEAD)
    {
        if(!w.constrained && shape == KS_WAVE && useAffine<T, EV, SA>(w) && w.maxPatch <= 64 && w.kind != EK_MIXED)
        {
            if(w.kind == EK_TET) launchNmAffine<T, EK_TET>(w, stream);
            else launchNmAffine<T, EK_HEX>(w, stream);
            return hipGetLastError();
        }
    }
    if(!w.constrained && shape == KS_WAVE && laneFits<T, EV, SA>(w, SM, lane))
    {
        if(w.kind == EK_TET) return launchLane<T, EV, SA, SM, EK_TET>(w, c, lane, moveCoeff, stream);
        return launchLane<T, EV, SA, SM, EK_HEX>(w, c, lane, moveCoeff, stream);
    }
    if(shape == KS_WAVE && w.maxPatch <= 64)
    {
        if(w.maxPatch <= 8)
            smoothWaveKernel<T, EV, SA, SM, 8><<<(w.count + 7) / 8, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, (T)moveCoeff);
        else if(w.maxPatch <= 16)
            smoothWaveKernel<T, EV, SA, SM, 16><<<(w.count + 3) / 4, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, (T)moveCoeff);
        else if(w.maxPatch <= 32)
            smoothWaveKernel<T, EV, SA, SM, 32><<<(w.count + 1) / 2, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, (T)moveCoeff);
        else
            smoothWaveKernel<T, EV, SA, SM, 64><<<w.count, 64, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, (T)moveCoeff);
        return hipGetLastError();
    }
    smoothPlainKernel<T, EV, SA, SM><<<(w.count + 255) / 256, 256, 0, stream>>>(w.meshes, w.segs, w.nsegs, w.count, c, (T)moveCoeff);
    return hipGetLastError();
}

template<typename T, int EV, int SA>
hipError_t Launchers<T, EV, SA>::smooth(const VertexWork& w, const gmq::Crew& c, int smoother, int shape,
                                        LaneShape lane, double moveCoeff, hipStream_t stream)
{
    if(smoother == SM_NELDER_MEAD)
        return launchSmoothTyped<T, EV, SA, SM_NELDER_MEAD>(w, c, shape, lane, moveCoeff, stream);
    if(smoother == SM_GRADIENT_DESCENT)
        return launchSmoothTyped<T, EV, SA, SM_GRADIENT_DESCENT>(w, c, shape, lane, moveCoeff, stream);
    if(smoother == SM_SPAWN_SEARCH)
        return launchSmoothTyped<T, EV, SA, SM_SPAWN_SEARCH>(w, c, shape, lane, moveCoeff, stream);
    return hipErrorInvalidValue;
}

template<typename T, int EV, int SA>
hipError_t Launchers<T, EV, SA>::getmeElements(const DevMesh* meshes, BlockWork tets, BlockWork hexs,
                                               BlockWork pris, const gmq::Crew& c, bool fast, hipStream_t stream)
{
    if(tets.count)
    {
        if(sizeof(T) == 8 && EV == 0 && SA == 0 && fast)
            getmeTetKernel<T, EV, SA, true><<<tets.count, ELEM_BLOCK, 0, stream>>>(meshes, tets.spans, tets.nspans, c);
        else
            getmeTetKernel<T, EV, SA><<<tets.count, ELEM_BLOCK, 0, stream>>>(meshes, tets.spans, tets.nspans, c);
    }
    if(hexs.count)
        getmeHexKernel<T, EV, SA><<<hexs.count, ELEM_BLOCK, 0, stream>>>(meshes, hexs.spans, hexs.nspans, c);
    if(pris.count)
        getmePriKernel<T, EV, SA><<<pris.count, ELEM_BLOCK, 0, stream>>>(meshes, pris.spans, pris.nspans, c);
    return hipGetLastError();
}

template<typename T, int EV, int SA, int EKIND, bool FASTQ = false>
static void launchGetmeWave(const VertexWork& w, const gmq::Crew& c, hipStream_t stream)
{
    const int threads = 64 * GETME_WAVES;
    if(w.maxPatch <= 8)
        getmeWaveKernel<T, EV, SA, 8, EKIND, FASTQ><<<(w.count + threads / 8 - 1) / (threads / 8), threads, 0, stream>>>(
            w.meshes, w.segs, w.nsegs, w.count, c);
    else if(w.maxPatch <= 16)
        getmeWaveKernel<T, EV, SA, 16, EKIND, FASTQ><<<(w.count + threads / 16 - 1) / (threads / 16), threads, 0, stream>>>(
            w.meshes, w.segs, w.nsegs, w.count, c);
    else if(w.maxPatch <= 32)
        getmeWaveKernel<T, EV, SA, 32, EKIND, FASTQ><<<(w.count + threads / 32 - 1) / (threads / 32), threads, 0, stream>>>(
            w.meshes, w.segs, w.nsegs, w.count, c);
    else
        getmeWaveKernel<T, EV, SA, 64, EKIND, FASTQ><<<(w.count + threads / 64 - 1) / (threads / 64), threads, 0, stream>>>(
            w.meshes, w.segs, w.nsegs, w.count, c);
}

template<typename T, int EV, int SA>
hipError_t Launchers<T, EV, SA>::getmeVertices(const VertexWork& w, const gmq::Crew& c, int shape, LaneShape lane,
                                               hipStream_t stream)
{
    if(w.count == 0)
        return hipSuccess;
    if(shape == KS_WAVE && w.maxPatch <= 64)
    {
        if constexpr(sizeof(T) == 8 && EV == 0 && SA == 0)
            if(w.kind == EK_TET && w.polish)
            {
                launchGetmeWave<T, EV, SA, EK_TET, true>(w, c, stream);
                return hipGetLastError();
            }
        if(w.kind == EK_TET) launchGetmeWave<T, EV, SA, EK_TET>(w, c, stream);
        else if(w.kind == EK_HEX) launchGetmeWave<T, EV, SA, EK_HEX>(w, c, stream);
        else launchGetmeWave<T, EV, SA, EK_MIXED>(w, c, stream);
        return hipGetLastError();
    }
    const int ppg = lane.patchesPerGroup >= 1 && lane.patchesPerGroup <= 8 ? lane.patchesPerGroup : 4;
    const int threads = 64 * ppg;
    getmeVertexKernel<T, EV, SA><<<(w.count + threads - 1) / threads, threads, 0, stream>>>(w.meshes, w.segs,
                                                                                            w.nsegs, w.count, c);
    return hipGetLastError();
}

template<int EV, int SA>
hipError_t Evaluators<EV, SA>::evaluateF64(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                           uint32_t nmeshes, const gmq::Crew& c, bool hexes, hipStream_t stream)
{
    if(blocks.count == 0 || nmeshes == 0)
        return hipSuccess;
    if(hexes)
        evalKernel<EV, SA, double4, true><<<blocks.count, EVAL_THREADS, 0, stream>>>(meshes, blocks.spans,
                                                                                    blocks.nspans, c);
    else
        evalKernel<EV, SA, double4, false><<<blocks.count, EVAL_THREADS, 0, stream>>>(meshes, blocks.spans,
                                                                                     blocks.nspans, c);
    evalFinalKernel<<<nmeshes, EVAL_THREADS, 0, stream>>>(meshes, meshList);
    return hipGetLastError();
}

template<int EV, int SA>
hipError_t Evaluators<EV, SA>::evaluateF32(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                           uint32_t nmeshes, const gmq::Crew& c, bool hexes, hipStream_t stream)
{
    if(blocks.count == 0 || nmeshes == 0)
        return hipSuccess;
    if(hexes)
        evalKernel<EV, SA, float4, true><<<blocks.count, EVAL_THREADS, 0, stream>>>(meshes, blocks.spans,
                                                                                   blocks.nspans, c);
    else
        evalKernel<EV, SA, float4, false><<<blocks.count, EVAL_THREADS, 0, stream>>>(meshes, blocks.spans,
                                                                                    blocks.nspans, c);
    evalFinalKernel<<<nmeshes, EVAL_THREADS, 0, stream>>>(meshes, meshList);
    return hipGetLastError();
}

} // namespace gmd

#endif // GM_KERNELS_H
