// gm_smoothers.h -- the per-vertex patch optimisers of the smoothing path,
// written against an abstract patch-quality functor so that the same code
// drives both kernel shapes:
//   * one vertex per lane (the plain kernel: the functor loops over the
//     patch's elements), and
//   * one vertex per group of lanes (the wavefront kernel: every lane of the
//     group holds one element of the patch, the functor reduces across lanes
//     and all lanes of the group run the optimiser's scalar logic in lock-step).
// The functor `f(p)` returns the patch quality with the vertex moved to p and
// counts the evaluation.  Every arithmetic step keeps the reference's
// operation order (cited per function) so the FP64 instantiation is
// bit-comparable with an FP64 CPU build of the same order.
#ifndef GM_SMOOTHERS_H
#define GM_SMOOTHERS_H

#include "gm_quality.h"

namespace gms
{
using gmq::v3;
using gmq::mk;

template<typename T> struct v4
{
    v3<T> p;
    T w;
};

template<typename T> GMQ void swap4(v4<T>& a, v4<T>& b) { v4<T> t = a; a = b; b = t; }

// Constraint snapping of a trial position (MeshTopo::snapToBoundary,
// DataStructures/Mesh.h:256-267): the identity for a free vertex.  The
// optimisers snap exactly where the reference does ("if constrained, pos =
// snap(pos)"), so an identity snap leaves every bit of a free vertex's run.
template<typename T> struct NoSnap
{
    GMQ_MEMBER v3<T> operator()(v3<T> p) const { return p; }
};

// NelderMeadSmoother::smoothVertices (Smoothers/VertexWise/NelderMeadSmoother.cpp:100-261),
// parameters :17-23.  Returns the new position of a free vertex at `pos`.
template<typename T, class F, class S = NoSnap<T>> GMQ v3<T> nelderMead(v3<T> pos, T localSize, F& f, const S& snap = S())
{
    const T valueConvergence = T(0.000100);
    const int securityCycleCount = 12;
    const T nodeShift = localSize * T(1.0 / 24.0);

    v4<T> vo;
    vo.p = pos;
    vo.w = f(pos);
    v4<T> s[4];
    s[0].p = pos + mk<T>(nodeShift, T(0), T(0)); s[0].w = T(0);
    s[1].p = pos + mk<T>(T(0), nodeShift, T(0)); s[1].w = T(0);
    s[2].p = pos + mk<T>(T(0), T(0), nodeShift); s[2].w = T(0);
    s[3] = vo;

    int cycle = 0;
    bool reset = false, terminated = false;
    while(!terminated)
    {
        for(int p = 0; p < 3; ++p)
        {
            s[p].p = snap(s[p].p);
            s[p].w = f(s[p].p);
        }
        if(s[0].w > s[1].w) swap4(s[0], s[1]);
        if(s[1].w > s[2].w) swap4(s[1], s[2]);
        if(s[2].w > s[3].w) swap4(s[2], s[3]);
        if(s[0].w > s[1].w) swap4(s[0], s[1]);
        if(s[1].w > s[2].w) swap4(s[1], s[2]);
        if(s[0].w > s[1].w) swap4(s[0], s[1]);

        for(; cycle < securityCycleCount; ++cycle)
        {
            v3<T> c = T(1 / 3.0) * (s[1].p + s[2].p + s[3].p);
            v3<T> x = snap(c + T(1) * (c - s[0].p));       // reflect, alpha = 1
            T fr = f(x);
            T fv = fr;
            v3<T> xr = x;
            if(s[3].w < fr)
            {
                x = snap(c + T(2) * (x - c));                // expand, gamma = 2
                T fe = f(x);
                fv = fe;
                if(fe <= fr) { x = xr; fv = fr; }
            }
            else if(s[1].w >= fr)
            {
                if(fr > s[0].w) x = snap(c + T(0.5) * (xr - c));  // outside contraction, beta = 0.5
                else x = snap(c + T(0.5) * (s[0].p - c));         // inside contraction
                fv = f(x);
            }
            v4<T> vtx;
            vtx.p = x;
            vtx.w = fv;
            if(vtx.w > s[3].w) swap4(s[3], vtx);
            if(vtx.w > s[2].w) swap4(s[2], vtx);
            if(vtx.w > s[1].w) swap4(s[1], vtx);
            if(vtx.w > s[0].w) swap4(s[0], vtx);
            if((s[3].w - s[1].w) < valueConvergence) { terminated = true; break; }
        }
        if(terminated || (cycle >= securityCycleCount && reset))
            break;
        s[0].p = vo.p - mk<T>(nodeShift, T(0), T(0)); s[0].w = vo.w - T(0);
        s[1].p = vo.p - mk<T>(T(0), nodeShift, T(0)); s[1].w = vo.w - T(0);
        s[2].p = vo.p - mk<T>(T(0), T(0), nodeShift); s[2].w = vo.w - T(0);
        s[3] = vo;
        reset = true;
        cycle = 0;
    }
    return snap(s[3].p);
}

// GradientDescentSmoother::smoothVertices (GradientDescentSmoother.cpp:73-202),
// parameters :14-15.
template<typename T, class F, class S = NoSnap<T>> GMQ v3<T> gradientDescent(v3<T> pos, T localSize, F& f,
                                                                            const S& snap = S())
{
    const int securityCycleCount = 5;
    T nodeShift = localSize * T(1.0 / 75.0);
    const T originalNodeShift = nodeShift;
    for(int c = 0; c < securityCycleCount; ++c)
    {
        T q0 = f(snap(pos + mk<T>(-nodeShift, T(0), T(0))));
        T q1 = f(snap(pos + mk<T>(nodeShift, T(0), T(0))));
        T q2 = f(snap(pos + mk<T>(T(0), -nodeShift, T(0))));
        T q3 = f(snap(pos + mk<T>(T(0), nodeShift, T(0))));
        T q4 = f(snap(pos + mk<T>(T(0), T(0), -nodeShift)));
        T q5 = f(snap(pos + mk<T>(T(0), T(0), nodeShift)));
        v3<T> g = mk<T>(q1 - q0, q3 - q2, q5 - q4);
        T gn = gmq::tsqrt(gmq::dot(g, g));
        if(gn == T(0))
            break;
        v3<T> shift = g * (nodeShift / gn);
        // Proposition 0 is the answer when no quality beats -inf (strict >).
        T bestQ = -__builtin_huge_val();
        T bestOff = T(-0.25);
        v3<T> bestP = snap(pos + shift * T(-0.25));
#pragma unroll 1
        for(int p = 0; p < 8; ++p)
        {
            const T off = p == 0 ? T(-0.25) : p == 1 ? T(0.00) : p == 2 ? T(0.10) : p == 3 ? T(0.20)
                        : p == 4 ? T(0.40) : p == 5 ? T(0.80) : p == 6 ? T(1.20) : T(1.60);
            v3<T> prop = snap(pos + shift * off);
            T q = f(prop);
            if(q > bestQ) { bestQ = q; bestOff = off; bestP = prop; }
        }
        pos = bestP;
        nodeShift *= gmq::tabs(bestOff);
        if(nodeShift < originalNodeShift / T(10))
            break;
    }
    return pos;
}

// SpawnSearchSmoother offsets (SpawnSearchSmoother.cpp:42-54): the 4^3
// lattice {-1.5,-0.5,0.5,1.5}^3, x fastest, with entry 0 replaced by the origin.
template<typename T> GMQ v3<T> spawnOffset(int p)
{
    if(p == 0) return mk<T>(T(0), T(0), T(0));
    return mk<T>((T)(p & 3) - T(1.5), (T)((p >> 2) & 3) - T(1.5), (T)(p >> 4) - T(1.5));
}

// SpawnSearchSmoother::smoothVertices (SpawnSearchSmoother.cpp:161-226).
template<typename T, class F, class S = NoSnap<T>> GMQ v3<T> spawnSearch(v3<T> pos, T localSize, T moveCoeff, F& f,
                                                                        const S& snap = S())
{
    T scale = moveCoeff * localSize;
    for(int iter = 0; iter < 2; ++iter)
    {
        T bestQ = -__builtin_huge_val();
        v3<T> bestP = snap(pos + spawnOffset<T>(0) * scale);
#pragma unroll 1
        for(int p = 0; p < 64; ++p)
        {
            v3<T> prop = snap(pos + spawnOffset<T>(p) * scale);
            T q = f(prop);
            if(q > bestQ) { bestQ = q; bestP = prop; }
        }
        pos = bestP;
        scale /= T(3);
    }
    return pos;
}

} // namespace gms

#endif // GM_SMOOTHERS_H
