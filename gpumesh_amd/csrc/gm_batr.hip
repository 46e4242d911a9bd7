// gm_batr.hip -- the BATR topologist (Topologists/BatrTopologist.cpp) on the
// host, between the device smoothing passes of a global pass.
//
// Its operations are greedy and sequential: each accepted swap, split or
// merge changes the patches the next candidate is judged on, and the
// reference visits candidates in the order of its per-vertex lists, which the
// operations themselves reorder (removal moves the last entry into the gap).
// The same results therefore need the same visiting order and the same list
// edits, kept here on per-vertex arrays; element qualities and edge lengths
// come from the routines the device kernels use (gm_quality.h, FP64), so a
// restructured mesh matches the reference's bit for bit.  The greedy choices
// also hinge on last-bit differences between nearly equal qualities
// (congruent configurations tie up to rounding), so the transcendentals are
// the C library's, the reference's libm, not the project's (the device's).
//
// Vertices are free or fixed.  For those two kinds the reference's boundary
// algebra (AbstractBoundary.cpp:39-66, VertexConstraint.cpp:65-150,
// VolumeConstraint.cpp:52-60) reduces to anchors: a fixed vertex holds its own
// anchor point, a free one none; merging keeps the anchor of either side
// (two different anchors cannot merge); a split vertex is free.  Its
// cureBoundaries step (:1486-1712) acts only on face-constrained vertices and
// on tets whose three constrained vertices split to a constrained support,
// which free and fixed vertices never produce.
#include "gm_batr.h"

#include <algorithm>
#include <array>
#include <cmath>

namespace gmb
{
namespace
{

using P3 = gmq::v3<double>;
using Tet = std::array<uint32_t, 4>;
using List = std::vector<uint32_t>;

const int32_t FREE = -1;

bool has(const List& l, uint32_t id) { return std::find(l.begin(), l.end(), id) != l.end(); }

// Removes the first `id`, moving the last entry into its place.
bool dropOne(List& l, uint32_t id)
{
    for(size_t i = 0; i < l.size(); ++i)
        if(l[i] == id)
        {
            l[i] = l.back();
            l.pop_back();
            return true;
        }
    return false;
}

// Removes every entry listed in `ids` with the same rule, scanning forward
// (an entry moved into a gap is examined next).
void dropAll(List& l, const List& ids)
{
    size_t i = 0;
    while(i < l.size())
    {
        if(has(ids, l[i]))
        {
            l[i] = l.back();
            l.pop_back();
        }
        else
            ++i;
    }
}

void addOnce(List& l, uint32_t id)
{
    if(!has(l, id)) l.push_back(id);
}

void replaceFirst(List& l, uint32_t from, uint32_t to)
{
    for(uint32_t& x : l)
        if(x == from)
        {
            x = to;
            return;
        }
}

// The corner of `t` equal to `from` becomes `to` (the first one).
void renameCorner(Tet& t, uint32_t from, uint32_t to)
{
    for(uint32_t& x : t)
        if(x == from)
        {
            x = to;
            return;
        }
}

// Ring triangulations (BatrTopologist.cpp:36-60): for each ring size, the
// triangulations of the ring polygon, each with the number of its distinct
// rotations.  The order is the reference's: ties keep the first.
struct Triangulation
{
    uint32_t rotations;
    std::vector<std::array<uint32_t, 3>> tris;
};
const std::vector<std::vector<Triangulation>>& triangulations()
{
    static const std::vector<std::vector<Triangulation>> table = {
        {}, {}, {},
        {{1, {{0, 1, 2}}}},
        {{2, {{0, 1, 2}, {0, 2, 3}}}},
        {{5, {{0, 1, 3}, {0, 3, 4}, {1, 2, 3}}}},
        {{6, {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5}}},
         {3, {{0, 1, 2}, {0, 2, 3}, {0, 3, 5}, {3, 4, 5}}},
         {3, {{0, 1, 3}, {1, 2, 3}, {0, 3, 4}, {0, 4, 5}}},
         {2, {{0, 1, 2}, {0, 2, 4}, {2, 3, 4}, {0, 4, 5}}}},
        {{7, {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 5}, {0, 5, 6}}},
         {7, {{0, 1, 2}, {0, 2, 3}, {0, 3, 6}, {3, 4, 6}, {4, 5, 6}}},
         {7, {{0, 1, 2}, {0, 2, 3}, {0, 3, 4}, {0, 4, 6}, {4, 5, 6}}},
         {7, {{0, 1, 3}, {1, 2, 3}, {0, 3, 4}, {0, 4, 5}, {0, 5, 6}}},
         {7, {{0, 1, 2}, {0, 2, 4}, {2, 3, 4}, {0, 4, 5}, {0, 5, 6}}},
         {7, {{0, 1, 2}, {0, 2, 3}, {0, 3, 5}, {3, 4, 5}, {0, 5, 6}}}}};
    return table;
}

// The face opposite each corner, wound to face inwards (MeshTet::tris,
// DataStructures/Mesh.cpp:30-35).
const int FACE[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

template<int EV, int SA> double tetQ(const gmq::Crew& c, const P3* vp)
{
    return gmq::tetQuality<double, EV, SA>(c, vp);
}

// Edge length in the crew's metric: MetricFreeMeasurer::riemannianDistance
// (MetricFreeMeasurer.cpp:27-34) for the uniform sampler, the metric-wise
// walk (MetricWiseMeasurer.cpp:65-116) otherwise.
template<int SA> double edgeLength(const gmq::Crew& c, P3 a, P3 b)
{
    if(SA == gmq::SA_UNIFORM)
    {
        const P3 d = a - b;
        return std::sqrt(gmq::dot(d, d)) * c.scaling;
    }
    return gmq::riemannianDistanceWise<double, SA>(c, a, b);
}

struct Work
{
    const Settings& s;
    double (*qualityFn)(const gmq::Crew&, const P3*) = nullptr;
    double (*lengthFn)(const gmq::Crew&, P3, P3) = nullptr;

    std::vector<P3> x;
    std::vector<Tet> tet;
    std::vector<List> tetsAt;    // the patch of each vertex
    std::vector<List> nbrsAt;    // its edge neighbours
    std::vector<int32_t> anchor; // FREE, or an index into anchorAt
    std::vector<P3> anchorAt;

    // Tet-membership stamps: a list's tets marked with the current
    // generation answer "is e in that list" in O(1).
    std::vector<uint32_t> stamp;
    uint32_t gen = 0;
    // Qualities of tet slots while no vertex moves (the swap operations);
    // a slot's entry is dropped whenever the slot is rewritten.
    std::vector<double> slotQ;
    std::vector<uint8_t> slotQOk;

    explicit Work(const Settings& st) : s(st) {}

    double quality(const Tet& t) const
    {
        const P3 vp[4] = {x[t[0]], x[t[1]], x[t[2]], x[t[3]]};
        return qualityFn(s.crew, vp);
    }
    double slotQuality(uint32_t e)
    {
        if(e >= slotQOk.size())
        {
            slotQOk.resize(tet.size(), 0);
            slotQ.resize(tet.size());
        }
        if(!slotQOk[e])
        {
            slotQ[e] = quality(tet[e]);
            slotQOk[e] = 1;
        }
        return slotQ[e];
    }
    void rewritten(uint32_t e)
    {
        if(e < slotQOk.size()) slotQOk[e] = 0;
    }
    void forgetQualities() { std::fill(slotQOk.begin(), slotQOk.end(), 0); }

    uint32_t nextGeneration(uint32_t step = 1)
    {
        if(stamp.size() < tet.size()) stamp.resize(tet.size(), 0);
        if(gen > 0xFFFFFFF0u)
        {
            std::fill(stamp.begin(), stamp.end(), 0);
            gen = 0;
        }
        gen += step;
        return gen;
    }
    void mark(const List& l, uint32_t g)
    {
        for(uint32_t e : l) stamp[e] = g;
    }
    bool marked(uint32_t e, uint32_t g) const { return e < stamp.size() && stamp[e] == g; }
    double length(uint32_t a, uint32_t b) const { return lengthFn(s.crew, x[a], x[b]); }

    // AbstractMeasurer::tetEuclideanVolume (AbstractMeasurer.cpp:75-84), times 6.
    double signedVolume6(const Tet& t) const
    {
        const P3 p3 = x[t[3]];
        return gmq::det(gmq::mcols<double>(p3 - x[t[0]], p3 - x[t[1]], p3 - x[t[2]]));
    }

    // The neighbours of v from its patch, ascending.
    void rebuildNeighbours(uint32_t v)
    {
        List n;
        for(uint32_t e : tetsAt[v])
            for(uint32_t c : tet[e])
                if(c != v) n.push_back(c);
        sortUnique(n);
        nbrsAt[v] = n;
    }

    // Ascending, duplicates dropped (short lists: insertion sort).
    static void sortUnique(List& l)
    {
        if(l.size() > 24)
        {
            std::sort(l.begin(), l.end());
            l.erase(std::unique(l.begin(), l.end()), l.end());
            return;
        }
        for(size_t i = 1; i < l.size(); ++i)
        {
            const uint32_t k = l[i];
            size_t j = i;
            while(j > 0 && l[j - 1] > k)
            {
                l[j] = l[j - 1];
                --j;
            }
            l[j] = k;
        }
        l.erase(std::unique(l.begin(), l.end()), l.end());
    }

    // The tets around edge a-b, in a's patch order, and their other
    // vertices, ascending (BatrTopologist::findRing, :1359-1393).
    void edgeRing(uint32_t a, uint32_t b, List& around, List& ring)
    {
        const uint32_t g = nextGeneration();
        mark(tetsAt[b], g);
        for(uint32_t e : tetsAt[a])
            if(marked(e, g)) around.push_back(e);
        for(uint32_t e : around)
            for(uint32_t c : tet[e])
                if(c != a && c != b) ring.push_back(c);
        sortUnique(ring);
    }

    // v's tets outside `around`, in patch order.
    void tetsOnlyAt(uint32_t v, const List& around, List& out)
    {
        const uint32_t g = nextGeneration();
        mark(around, g);
        for(uint32_t e : tetsAt[v])
            if(!marked(e, g)) out.push_back(e);
    }

    // Drops dead tets, moving the survivors down and renaming them in the
    // patches of their corners (BatrTopologist::trimTets, :1457-1484).
    void compactTets(const std::vector<bool>& alive)
    {
        size_t to = 0;
        const size_t n = tet.size();
        for(size_t e = 0; e < n; ++e)
        {
            if(!alive[e]) continue;
            if(to != e)
            {
                const Tet t = tet[e];
                for(uint32_t c : t) replaceFirst(tetsAt[c], (uint32_t)e, (uint32_t)to);
                tet[to] = t;
            }
            ++to;
        }
        tet.resize(to);
        forgetQualities();
    }

    // Drops dead vertices the same way (BatrTopologist::trimVerts, :1420-1455).
    void compactVertices(const std::vector<bool>& alive)
    {
        size_t to = 0;
        for(size_t v = 0; v < x.size(); ++v)
        {
            if(!alive[v]) continue;
            if(to != v)
            {
                for(uint32_t n : nbrsAt[v]) replaceFirst(nbrsAt[n], (uint32_t)v, (uint32_t)to);
                for(uint32_t e : tetsAt[v]) renameCorner(tet[e], (uint32_t)v, (uint32_t)to);
                x[to] = x[v];
                tetsAt[to] = tetsAt[v];
                nbrsAt[to] = nbrsAt[v];
                anchor[to] = anchor[v];
            }
            ++to;
        }
        x.resize(to);
        tetsAt.resize(to);
        nbrsAt.resize(to);
        anchor.resize(to);
    }
};

// Tet slots freed by an operation are reused last-freed first; otherwise a
// tet is appended.
uint32_t placeTet(Work& w, const Tet& t, List& freeTets, std::vector<bool>& alive)
{
    uint32_t id;
    if(freeTets.empty())
    {
        id = (uint32_t)w.tet.size();
        w.tet.push_back(t);
        alive.push_back(true);
    }
    else
    {
        id = freeTets.back();
        freeTets.pop_back();
        w.tet[id] = t;
        alive[id] = true;
    }
    w.rewritten(id);
    return id;
}

// 2-3 face swaps (BatrTopologist::faceSwapping, :635-836): two tets sharing
// a face become three around the segment joining their apexes when that
// raises the harmonic mean of their qualities.
int64_t swapFaces(Work& w)
{
    int64_t swaps = 0;
    std::vector<bool> pending(w.tet.size(), true);
    bool again = true;
    while(again)
    {
        again = false;
        for(size_t t = 0; t < w.tet.size(); ++t)
        {
            if(!pending[t]) continue;
            pending[t] = false;
            const Tet cur = w.tet[t];
            for(int f = 0; f < 4; ++f)
            {
                const uint32_t a = cur[FACE[f][0]], b = cur[FACE[f][1]], c = cur[FACE[f][2]];
                // Each interior face once: from the side where its ids rise
                // cyclically.
                if(!((a < b && b < c) || (b < c && c < a) || (c < a && a < b))) continue;
                uint32_t pair[2] = {0, 0};
                int shared = 0;
                const uint32_t g = w.nextGeneration(2);
                w.mark(w.tetsAt[b], g - 1);
                for(uint32_t e : w.tetsAt[c])
                    if(w.stamp[e] == g - 1) w.stamp[e] = g;
                for(uint32_t e : w.tetsAt[a])
                    if(w.marked(e, g))
                    {
                        if(shared < 2) pair[shared] = e;
                        ++shared;
                    }
                if(shared != 2) continue;

                const uint32_t top = cur[f];
                const uint32_t nt = (pair[0] != (uint32_t)t) ? pair[0] : pair[1];
                const Tet other = w.tet[nt];
                uint32_t bottom = other[0];
                for(int k = 1; k < 4; ++k)
                    if(other[k] != a && other[k] != b && other[k] != c)
                    {
                        bottom = other[k];
                        break;
                    }

                // The segment top-bottom must cross the face.
                const P3 o = w.x[top];
                const P3 pa = w.x[a] - o, pb = w.x[b] - o, pc = w.x[c] - o;
                const P3 d = w.x[bottom] - o;
                if(gmq::dot(gmq::cross(pa, pb), d) < 0.0 || gmq::dot(gmq::cross(pb, pc), d) < 0.0 ||
                   gmq::dot(gmq::cross(pc, pa), d) < 0.0)
                    continue;

                const double before = 2.0 / (1.0 / w.slotQuality((uint32_t)t) + 1.0 / w.slotQuality(nt));
                const Tet t0 = {top, a, b, bottom}, t1 = {top, b, c, bottom}, t2 = {top, c, a, bottom};
                const double q0 = w.quality(t0), q1 = w.quality(t1), q2 = w.quality(t2);
                const double after = 3.0 / (1.0 / q0 + 1.0 / q1 + 1.0 / q2);
                if(after <= before) continue;

                w.tet[t] = t0;
                w.tet[nt] = t1;
                const uint32_t lt = (uint32_t)w.tet.size();
                w.tet.push_back(t2);
                w.rewritten((uint32_t)t);
                w.rewritten(nt);
                w.rewritten(lt);
                w.tetsAt[top].push_back(nt);
                w.tetsAt[top].push_back(lt);
                w.nbrsAt[top].push_back(bottom);
                w.tetsAt[bottom].push_back((uint32_t)t);
                w.tetsAt[bottom].push_back(lt);
                w.nbrsAt[bottom].push_back(top);
                replaceFirst(w.tetsAt[a], nt, lt);
                replaceFirst(w.tetsAt[c], (uint32_t)t, lt);
                pending.push_back(true);
                pending[nt] = true;
                pending[t] = true;
                again = true;
                ++swaps;
                break;
            }
        }
    }
    return swaps;
}

// Orders the vertices around edge v-n into a chain by walking its tets from
// the lowest id; at an open end the chain so far is reversed and the walk
// resumes on the other side (BatrTopologist.cpp:889-962).  False when a
// second end is met before every vertex is placed.
bool chainRing(const Work& w, uint32_t v, uint32_t n, const List& around, List& ring)
{
    List pool = around;
    int ends = 0;
    size_t cur = 0;
    while(cur + 1 < ring.size())
    {
        bool found = false;
        for(size_t i = 0; i < pool.size() && !found; ++i)
        {
            const Tet t = w.tet[pool[i]];
            for(int j = 0; j < 4 && !found; ++j)
            {
                if(t[j] != ring[cur]) continue;
                for(int k = 0; k < 4 && !found; ++k)
                {
                    if(k == j || t[k] == v || t[k] == n) continue;
                    for(size_t s = cur + 1; s < ring.size() && !found; ++s)
                        if(ring[s] == t[k])
                        {
                            pool[i] = pool.back();
                            pool.pop_back();
                            std::swap(ring[s], ring[cur + 1]);
                            found = true;
                        }
                }
            }
        }
        if(found)
        {
            ++cur;
            continue;
        }
        if(++ends >= 2) return false;
        std::reverse(ring.begin(), ring.begin() + cur + 1);
    }
    return true;
}

// Edge swaps (BatrTopologist::edgeSwapping, :838-1278): the tets around an
// edge are replaced by the triangulation of its ring (joined to both ends)
// whose worst tet beats the current worst.  Returns 0 when a ring cannot be
// closed, as the reference does.
int64_t swapEdges(Work& w)
{
    const auto& table = triangulations();
    std::vector<bool> pending(w.x.size(), true);
    std::vector<bool> alive(w.tet.size(), true);
    List freeTets;
    int64_t swaps = 0;
    bool again = true;
    while(again)
    {
        again = false;
        for(uint32_t v = 0; v < (uint32_t)w.x.size(); ++v)
        {
            if(!pending[v]) continue;
            pending[v] = false;
            for(size_t i = 0; i < w.nbrsAt[v].size(); ++i)
            {
                const uint32_t n = w.nbrsAt[v][i];
                if(n < v) continue;
                List around, ring;
                w.edgeRing(v, n, around, ring);
                if(!chainRing(w, v, n, around, ring))
                {
                    w.compactTets(alive);
                    return 0;
                }
                const size_t R = ring.size();
                if(R < 3 || R >= table.size()) continue;

                // Wind the ring counter-clockwise around v -> n.
                const P3 r0 = w.x[ring[0]];
                P3 area = gmq::mk<double>(0.0, 0.0, 0.0);
                for(size_t k = 2; k < R; ++k) area = area + gmq::cross(w.x[ring[k - 1]] - r0, w.x[ring[k]] - r0);
                const P3 pv = w.x[v];
                const P3 axis = w.x[n] - pv;
                if(gmq::dot(area, axis) < 0.0) std::reverse(ring.begin(), ring.end());

                double worst = INFINITY;
                for(uint32_t e : around) worst = gmq::tmin(worst, w.slotQuality(e));

                // Every ring edge must turn the same way around the axis.
                bool convex = true;
                for(size_t k = 1; k <= R; ++k)
                {
                    const P3 nrm = gmq::cross(w.x[ring[k - 1]] - pv, w.x[ring[k % R]] - pv);
                    if(gmq::dot(axis, nrm) < 0.0)
                    {
                        convex = false;
                        break;
                    }
                }
                if(!convex) continue;

                const std::vector<Triangulation>& options = table[R];
                int best = -1;
                uint32_t bestRot = 0;
                for(size_t o = 0; o < options.size(); ++o)
                    for(uint32_t rot = 0; rot < options[o].rotations; ++rot)
                    {
                        double low = INFINITY;
                        for(const auto& tri : options[o].tris)
                        {
                            const uint32_t a = ring[(tri[0] + rot) % R], b = ring[(tri[1] + rot) % R],
                                           c = ring[(tri[2] + rot) % R];
                            const double qv = w.quality(Tet{b, a, c, v});
                            low = gmq::tmin(low, qv);
                            if(qv < worst) break;
                            const double qn = w.quality(Tet{a, b, c, n});
                            low = gmq::tmin(low, qn);
                            if(qn < worst) break;
                        }
                        if(low > worst)
                        {
                            best = (int)o;
                            bestRot = rot;
                            worst = low;
                        }
                    }
                if(best < 0) continue;
                ++swaps;

                dropAll(w.tetsAt[v], around);
                dropAll(w.tetsAt[n], around);
                dropOne(w.nbrsAt[v], n);
                dropOne(w.nbrsAt[n], v);
                for(uint32_t r : ring) dropAll(w.tetsAt[r], around);
                for(uint32_t e : around)
                {
                    alive[e] = false;
                    freeTets.push_back(e);
                }
                std::vector<Tet> made;
                for(const auto& tri : options[best].tris)
                {
                    const uint32_t a = ring[(tri[0] + bestRot) % R], b = ring[(tri[1] + bestRot) % R],
                                   c = ring[(tri[2] + bestRot) % R];
                    made.push_back(Tet{b, a, c, v});
                    made.push_back(Tet{a, b, c, n});
                    addOnce(w.nbrsAt[a], b);
                    addOnce(w.nbrsAt[a], c);
                    addOnce(w.nbrsAt[b], a);
                    addOnce(w.nbrsAt[b], c);
                    addOnce(w.nbrsAt[c], a);
                    addOnce(w.nbrsAt[c], b);
                }
                for(const Tet& t : made)
                {
                    const uint32_t id = placeTet(w, t, freeTets, alive);
                    for(uint32_t c : t) w.tetsAt[c].push_back(id);
                }
                again = true;
                for(uint32_t u : w.nbrsAt[v]) pending[u] = true;
                for(uint32_t u : w.nbrsAt[n]) pending[u] = true;
            }
        }
    }
    w.compactTets(alive);
    return swaps;
}

// Anchor of a merged vertex (AbstractBoundary::merge for free and fixed
// vertices): false when two different anchors meet.
bool mergedAnchor(int32_t a, int32_t b, int32_t& out)
{
    if(a == b) out = a;
    else if(a == FREE) out = b;
    else if(b == FREE) out = a;
    else return false;
    return true;
}
// Anchor of a split vertex (AbstractBoundary::split): free unless both ends
// hold the same one.
int32_t splitAnchor(int32_t a, int32_t b) { return a == b ? a : FREE; }

// Edge splits and merges (BatrTopologist::edgeSplitMerge, :137-633): per
// vertex, the edge furthest outside [minEdge, maxEdge] is merged (too short)
// or split at its middle (too long) when no affected tet falls below the
// current worst divided by how far outside the edge is.
void splitMerge(Work& w, int64_t& splits, int64_t& merges)
{
    const Settings& s = w.s;
    const double FLOOR = 1e-6;   // BatrTopologist::_minAcceptableGenQuality
    std::vector<bool> aliveV(w.x.size(), true), aliveT(w.tet.size(), true);
    std::vector<bool> pending(w.x.size(), true);
    List freeV, freeT;
    bool again = true;
    for(int sweep = 0; sweep < s.refinementSweeps && again; ++sweep)
    {
        again = false;
        for(uint32_t v = 0; v < (uint32_t)w.x.size(); ++v)
        {
            if(!aliveV[v]) continue;
            pending[v] = false;

            uint32_t n = 0;
            double len = INFINITY, urgency = 0.0;
            int32_t keep = FREE;
            bool picked = false;
            for(size_t i = 0; i < w.nbrsAt[v].size(); ++i)
            {
                const uint32_t cand = w.nbrsAt[v][i];
                if(v > cand && !pending[cand]) continue;
                const double d = w.length(v, cand);
                if(d < s.minEdge)
                {
                    const double u = s.minEdge / d;
                    int32_t a;
                    if(u > urgency && mergedAnchor(w.anchor[v], w.anchor[cand], a))
                    {
                        keep = a;
                        picked = true;
                        urgency = u;
                        len = d;
                        n = cand;
                    }
                }
                else if(d > s.maxEdge)
                {
                    const double u = d / s.maxEdge;
                    if(u > urgency)
                    {
                        keep = splitAnchor(w.anchor[v], w.anchor[cand]);
                        picked = true;
                        urgency = u;
                        len = d;
                        n = cand;
                    }
                }
            }
            if(!picked) continue;

            const P3 pv = w.x[v], pn = w.x[n];
            List around, ring, onlyN;
            w.edgeRing(v, n, around, ring);
            w.tetsOnlyAt(n, around, onlyN);
            P3 mid = (pv + pn) / 2.0;
            if(keep != FREE) mid = w.anchorAt[keep];

            if(len < s.minEdge)
            {
                // Merge n into v at the middle (or at the anchor).
                List onlyV;
                w.tetsOnlyAt(v, around, onlyV);
                List touched = onlyV;
                touched.insert(touched.end(), onlyN.begin(), onlyN.end());
                double floorQ = 1.0;
                for(uint32_t e : touched) floorQ = gmq::tmin(floorQ, w.quality(w.tet[e]));
                floorQ = gmq::tmax(floorQ / urgency, FLOOR);
                w.x[v] = mid;
                w.x[n] = mid;
                bool ok = true;
                for(uint32_t e : touched)
                    if(w.quality(w.tet[e]) <= floorQ)
                    {
                        ok = false;
                        break;
                    }
                if(!ok)
                {
                    w.x[v] = pv;
                    w.x[n] = pn;
                    continue;
                }
                freeV.push_back(n);
                aliveV[n] = false;
                for(uint32_t e : around)
                {
                    freeT.push_back(e);
                    aliveT[e] = false;
                }
                for(uint32_t e : onlyN) renameCorner(w.tet[e], n, v);
                for(uint32_t r : ring)
                {
                    dropAll(w.tetsAt[r], around);
                    if(w.tetsAt[r].empty())
                    {
                        aliveV[r] = false;
                        freeV.push_back(r);
                    }
                }
                w.anchor[v] = keep;
                w.tetsAt[v] = touched;
                w.rebuildNeighbours(v);
                const List nb = w.nbrsAt[v];
                for(uint32_t u : nb) w.rebuildNeighbours(u);
                ++merges;
                again = true;
                pending[v] = true;
                pending[n] = false;
                for(uint32_t u : w.nbrsAt[v]) pending[u] = true;
            }
            else if(len > s.maxEdge)
            {
                // Split v-n at a new vertex m.
                uint32_t m;
                if(freeV.empty())
                {
                    m = (uint32_t)w.x.size();
                    w.x.push_back(mid);
                    w.anchor.push_back(keep);
                    w.tetsAt.emplace_back();
                    w.nbrsAt.emplace_back();
                    aliveV.push_back(true);
                    pending.push_back(true);
                }
                else
                {
                    m = freeV.back();
                    freeV.pop_back();
                    w.x[m] = mid;
                    w.anchor[m] = keep;
                    w.tetsAt[m].clear();
                    w.nbrsAt[m].clear();
                    aliveV[m] = true;
                    pending[m] = true;
                }
                double floorQ = 1.0;
                for(uint32_t e : around) floorQ = gmq::tmin(floorQ, w.quality(w.tet[e]));
                floorQ = gmq::tmax(floorQ / urgency, FLOOR);
                // The halves of a tet of the ring: v's side (n -> m) and n's side (v -> m).
                auto halves = [&](const Tet& t, Tet& sv, Tet& sn) {
                    sv = t;
                    sn = t;
                    for(int k = 0; k < 4; ++k)
                    {
                        if(t[k] == v) sn[k] = m;
                        else if(t[k] == n) sv[k] = m;
                    }
                };
                bool ok = true;
                for(uint32_t e : around)
                {
                    Tet sv, sn;
                    halves(w.tet[e], sv, sn);
                    const double qv = w.quality(sv), qn = w.quality(sn);
                    if(qv <= floorQ || qn <= floorQ)
                    {
                        ok = false;
                        break;
                    }
                }
                if(!ok)
                {
                    pending[m] = false;
                    aliveV[m] = false;
                    freeV.push_back(m);
                    continue;
                }
                for(uint32_t e : around)
                {
                    const Tet t = w.tet[e];
                    for(uint32_t c : t) dropOne(w.tetsAt[c], e);
                    freeT.push_back(e);
                    aliveT[e] = false;
                    Tet sv, sn;
                    halves(t, sv, sn);
                    if(w.signedVolume6(sv) / 6.0 < 0.0) std::swap(sv[0], sv[1]);
                    if(w.signedVolume6(sn) / 6.0 < 0.0) std::swap(sn[0], sn[1]);
                    const uint32_t iv = placeTet(w, sv, freeT, aliveT);
                    const uint32_t in = placeTet(w, sn, freeT, aliveT);
                    for(uint32_t c : sv) w.tetsAt[c].push_back(iv);
                    for(uint32_t c : sn) w.tetsAt[c].push_back(in);
                }
                w.rebuildNeighbours(m);
                const List nb = w.nbrsAt[m];
                for(uint32_t u : nb)
                {
                    w.rebuildNeighbours(u);
                    pending[u] = true;
                }
                ++splits;
                again = true;
            }
        }
    }
    w.compactTets(aliveT);
    w.compactVertices(aliveV);
}

template<int EV> void pickSampler(Work& w, int sampler)
{
    switch(sampler)
    {
    case gmq::SA_SHELL:
        w.qualityFn = tetQ<EV, gmq::SA_SHELL>;
        w.lengthFn = edgeLength<gmq::SA_SHELL>;
        break;
    case gmq::SA_TEXTURE:
        w.qualityFn = tetQ<EV, gmq::SA_TEXTURE>;
        w.lengthFn = edgeLength<gmq::SA_TEXTURE>;
        break;
    case gmq::SA_ANALYTIC:
        w.qualityFn = tetQ<EV, gmq::SA_ANALYTIC>;
        w.lengthFn = edgeLength<gmq::SA_ANALYTIC>;
        break;
    default:
        w.qualityFn = tetQ<EV, gmq::SA_UNIFORM>;
        w.lengthFn = edgeLength<gmq::SA_UNIFORM>;
    }
}

} // namespace

bool restructure(const TetMesh& in, const Settings& s, TetMesh& out, Counts& counts, std::string& err)
{
    counts = Counts{};
    const size_t V = in.fixed.size();
    const size_t T = in.tets.size() / 4;
    if(in.pos.size() != 3 * V || in.tets.size() != 4 * T || in.eoff.size() != V + 1 || in.noff.size() != V + 1 ||
       in.eref.size() != in.eoff[V] || in.nbr.size() != in.noff[V])
    {
        err = "inconsistent mesh arrays";
        return false;
    }
    for(uint32_t c : in.tets)
        if(c >= V)
        {
            err = "tet vertex out of range";
            return false;
        }
    if(s.sampler == gmq::SA_TEXTURE && !s.crew.texels)
    {
        err = "texture sampler without texels";
        return false;
    }

#if !defined(__HIP_DEVICE_COMPILE__)
    // The C library's pow, exp and cos for this thread while it runs.
    struct LibcScope
    {
        bool old = gmq::hostLibc;
        LibcScope() { gmq::hostLibc = true; }
        ~LibcScope() { gmq::hostLibc = old; }
    } libc;
#endif

    Work w(s);
    if(s.evaluator == gmq::EV_METRIC_CONFORMITY) pickSampler<gmq::EV_METRIC_CONFORMITY>(w, s.sampler);
    else pickSampler<gmq::EV_MEAN_RATIO>(w, s.sampler);
    w.x.resize(V);
    w.anchor.assign(V, FREE);
    w.tetsAt.resize(V);
    w.nbrsAt.resize(V);
    for(size_t v = 0; v < V; ++v)
    {
        w.x[v] = gmq::mk<double>(in.pos[3 * v], in.pos[3 * v + 1], in.pos[3 * v + 2]);
        if(in.fixed[v])
        {
            w.anchor[v] = (int32_t)w.anchorAt.size();
            w.anchorAt.push_back(w.x[v]);
        }
        for(uint32_t k = in.eoff[v]; k < in.eoff[v + 1]; ++k) w.tetsAt[v].push_back(in.eref[k] >> 4);
        w.nbrsAt[v].assign(in.nbr.begin() + in.noff[v], in.nbr.begin() + in.noff[v + 1]);
    }
    w.tet.resize(T);
    for(size_t e = 0; e < T; ++e)
        for(int k = 0; k < 4; ++k) w.tet[e][k] = in.tets[4 * e + k];

    // BatrTopologist::restructureMesh (:76-128).
    int64_t lastOps = 0;
    while(true)
    {
        const int64_t fs = swapFaces(w);
        const int64_t es = swapEdges(w);
        int64_t sp = 0, mg = 0;
        splitMerge(w, sp, mg);
        counts.faceSwaps += fs;
        counts.edgeSwaps += es;
        counts.splits += sp;
        counts.merges += mg;
        const int64_t ops = fs + es + sp + mg;
        ++counts.passes;
        if(ops == 0 || ops == lastOps || counts.passes >= s.topoPasses) break;
        lastOps = ops;
    }

    const size_t V2 = w.x.size();
    out = TetMesh{};
    out.pos.resize(3 * V2);
    out.fixed.resize(V2);
    for(size_t v = 0; v < V2; ++v)
    {
        out.pos[3 * v] = w.x[v].x;
        out.pos[3 * v + 1] = w.x[v].y;
        out.pos[3 * v + 2] = w.x[v].z;
        out.fixed[v] = w.anchor[v] != FREE ? 1 : 0;
    }
    out.tets.resize(4 * w.tet.size());
    for(size_t e = 0; e < w.tet.size(); ++e)
        for(int k = 0; k < 4; ++k) out.tets[4 * e + k] = w.tet[e][k];
    return true;
}

} // namespace gmb
