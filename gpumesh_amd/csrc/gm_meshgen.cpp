// gm_meshgen.cpp -- structured unit-cube meshes with seeded jitter (the
// synthetic inputs of the smoothing workloads), generated natively so that
// benchmarks and serving streams can build hundreds of meshes per second.
//
// Grid numbering follows the reference's grid mesher (x fastest,
// Meshers/DebugMesher.cpp:173-181); hexahedra list v0..v3 counter-clockwise
// on the bottom face and v4..v7 on top (DebugMesher.cpp:285-293).  Kuhn
// tetrahedra are the 6 monotone lattice paths of a cell, oriented so that
// det(v3-v0, v3-v1, v3-v2) > 0 (MeanRatioEvaluator.cpp:34-50 convention).
// Interior vertices get a per-axis displacement U[-a h, a h] from
// SplitMix64(seed, 3 v + axis); elements whose orientation determinant falls
// below a fraction of the unjittered one have the jitter of their vertices
// halved until none remain.  Every floating-point step is written in a fixed
// order without contraction, so the output is identical on every host.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/gpumesh.h"

namespace
{

uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t seed, uint64_t counter)
{
    const uint64_t key = seed * 0xD1B54A32D192ED03ull + counter;
    return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

struct V3
{
    double x, y, z;
};

inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 neg(const V3& a) { return {-a.x, -a.y, -a.z}; }
inline V3 cross(const V3& b, const V3& c)
{
    return {b.y * c.z - b.z * c.y, b.z * c.x - b.x * c.z, b.x * c.y - b.y * c.x};
}
inline double dot(const V3& a, const V3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 at(const std::vector<double>& p, uint32_t v) { return {p[3 * v], p[3 * v + 1], p[3 * v + 2]}; }

double tetDet(const std::vector<double>& p, const uint32_t* t)
{
    const V3 p0 = at(p, t[0]), p1 = at(p, t[1]), p2 = at(p, t[2]), p3 = at(p, t[3]);
    return dot(sub(p3, p0), cross(sub(p3, p1), sub(p3, p2)));
}

// Smallest of the 8 corner-frame determinants of MeanRatioEvaluator::hexQuality
// (MeanRatioEvaluator.cpp:127-134).
double hexMinCornerDet(const std::vector<double>& p, const uint32_t* h)
{
    V3 q[8];
    for(int k = 0; k < 8; ++k) q[k] = at(p, h[k]);
    auto e = [&](int i, int j) { return sub(q[j], q[i]); };
    const V3 f[8][3] = {{e(0, 1), e(0, 4), neg(e(0, 3))}, {e(0, 1), e(1, 2), e(1, 5)},
                        {e(1, 2), e(2, 6), neg(e(2, 3))}, {e(0, 3), e(2, 3), e(3, 7)},
                        {e(0, 4), e(4, 5), e(4, 7)},      {e(1, 5), neg(e(5, 6)), e(4, 5)},
                        {e(2, 6), e(5, 6), e(6, 7)},      {e(3, 7), e(6, 7), neg(e(4, 7))}};
    double m = dot(f[0][0], cross(f[0][1], f[0][2]));
    for(int k = 1; k < 8; ++k)
    {
        const double d = dot(f[k][0], cross(f[k][1], f[k][2]));
        if(d < m) m = d;
    }
    return m;
}

} // namespace

extern "C" int gm_cube_mesh(int n, int kind, double jitter, uint64_t seed, double minDetFraction, double* pos,
                            int32_t* fixed, uint32_t* elems)
{
    if(n < 1 || n > 1000 || (kind != 0 && kind != 1) || !pos || !fixed || !elems || !(jitter >= 0.0))
        return GM_ERR_ARG;
    const uint64_t w = (uint64_t)n + 1;
    const uint64_t V = w * w * w;
    const uint64_t cells = (uint64_t)n * n * n;
    const int per = kind == 0 ? 4 : 8;
    const uint64_t E = kind == 0 ? 6 * cells : cells;
    const double h = 1.0 / n;

    std::vector<double> base(3 * V), disp(3 * V);
    const double amp = jitter * h;
    for(uint64_t k = 0, v = 0; k < w; ++k)
        for(uint64_t j = 0; j < w; ++j)
            for(uint64_t i = 0; i < w; ++i, ++v)
            {
                base[3 * v] = (double)i / (double)n;
                base[3 * v + 1] = (double)j / (double)n;
                base[3 * v + 2] = (double)k / (double)n;
                const bool bnd = i == 0 || j == 0 || k == 0 || i == (uint64_t)n || j == (uint64_t)n || k == (uint64_t)n;
                fixed[v] = bnd ? 1 : 0;
                for(int d = 0; d < 3; ++d)
                {
                    const double u = uniform01(seed, 3 * v + (uint64_t)d);
                    disp[3 * v + d] = bnd ? 0.0 : (2.0 * u - 1.0) * amp;
                }
            }

    // Cell corner c = bx + 2 by + 4 bz.
    uint64_t off[8];
    for(int c = 0; c < 8; ++c) off[c] = (uint64_t)(c & 1) + (uint64_t)((c >> 1) & 1) * w + (uint64_t)((c >> 2) & 1) * w * w;
    int kuhn[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};
    for(auto& t : kuhn)
    {
        // Orientation on the unit cell (integer coordinates: exact).
        long long c[4][3];
        for(int a = 0; a < 4; ++a) { c[a][0] = t[a] & 1; c[a][1] = (t[a] >> 1) & 1; c[a][2] = (t[a] >> 2) & 1; }
        long long a0[3], a1[3], a2[3];
        for(int d = 0; d < 3; ++d) { a0[d] = c[3][d] - c[0][d]; a1[d] = c[3][d] - c[1][d]; a2[d] = c[3][d] - c[2][d]; }
        const long long det = a0[0] * (a1[1] * a2[2] - a1[2] * a2[1]) - a0[1] * (a1[0] * a2[2] - a1[2] * a2[0]) +
                              a0[2] * (a1[0] * a2[1] - a1[1] * a2[0]);
        if(det < 0) { int s = t[1]; t[1] = t[2]; t[2] = s; }
    }
    const int hexCodes[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    uint64_t e = 0;
    for(uint64_t k = 0; k < (uint64_t)n; ++k)
        for(uint64_t j = 0; j < (uint64_t)n; ++j)
            for(uint64_t i = 0; i < (uint64_t)n; ++i)
            {
                const uint64_t cb = i + j * w + k * w * w;
                if(kind == 0)
                    for(int t = 0; t < 6; ++t, ++e)
                        for(int a = 0; a < 4; ++a) elems[4 * e + a] = (uint32_t)(cb + off[kuhn[t][a]]);
                else
                {
                    for(int a = 0; a < 8; ++a) elems[8 * e + a] = (uint32_t)(cb + off[hexCodes[a]]);
                    ++e;
                }
            }

    double refDet;
    if(kind == 0)
    {
        refDet = std::fabs(tetDet(base, elems));
        for(int t = 1; t < 6; ++t) refDet = std::fmin(refDet, std::fabs(tetDet(base, elems + 4 * t)));
    }
    else
        refDet = std::pow(h, 3.0);
    const double threshold = minDetFraction * refDet;

    std::vector<double> scale(V, 1.0), p(3 * V);
    std::vector<uint8_t> mark(V);
    for(int iter = 0;; ++iter)
    {
        for(uint64_t v = 0; v < V; ++v)
            for(int d = 0; d < 3; ++d) p[3 * v + d] = base[3 * v + d] + disp[3 * v + d] * scale[v];
        bool any = false;
        std::fill(mark.begin(), mark.end(), 0);
        for(uint64_t q = 0; q < E; ++q)
        {
            const uint32_t* el = elems + (size_t)per * q;
            const double d = kind == 0 ? tetDet(p, el) : hexMinCornerDet(p, el);
            if(d < threshold)
            {
                any = true;
                for(int a = 0; a < per; ++a) mark[el[a]] = 1;
            }
        }
        if(!any) break;
        if(iter == 63) return GM_ERR_ARG;
        for(uint64_t v = 0; v < V; ++v)
            if(mark[v]) scale[v] *= 0.5;
    }
    for(uint64_t q = 0; q < 3 * V; ++q) pos[q] = p[q];
    return GM_OK;
}
