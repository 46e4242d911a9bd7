// oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 CPU restatement of GpuMesh's vertex-relocation smoothing path,
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker.  Nothing in the product (gpumesh_amd/) links, loads or calls
// this file.  Each routine cites the reference file:line it follows; the code
// is written from those lines' behaviour, keeping the reference's operation
// order because last-ulp differences change optimiser branches.
//
// GLM (the reference's vector library, an un-vendored dependency with no
// pinned version) is replaced by the small subset below; its formulas follow
// GLM's published definitions (column-major matrices, cofactor determinant
// and inverse, length = sqrt(dot), normalize = v * (1/sqrt(dot)),
// mix = x + a(y - x)).  Parity against the real GLM build is unpinned.
//
// Transcendentals: libm mode 0 calls glibc (std::pow / std::exp) as the
// reference does; libm mode 1 substitutes the project's gm_libm.h routines,
// which the GPU kernels also use, so FP64 GPU results can be compared
// bit-for-bit.  The distance between the two modes is measured by the tests.
//
// What pins it (the reference ships no golden vectors): tests/test_oracle.py
// and tests/test_oracle_pins.py check element qualities against closed forms;
// every trial point and value of the Nelder-Mead, gradient-descent and
// spawn-search vertex steps, the GETMe element transform and vertex step, and
// two whole passes of each, against versions re-derived from the reference
// sources and driven by those closed forms; colourings against cases derived
// by hand; the Keast rule against exact integrals of a trilinear metric; the
// metric-wise walk against the error band of its 1 % rule.  Parity unpinned:
// the GLM boundary (formula order of its matrix product, determinant, inverse
// and mix; version unknown), which moves results only in the last bits.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <set>
#include <thread>
#include <vector>

#include "../gpumesh_amd/csrc/gm_libm.h"

namespace orc
{

// ------------------------------------------------------------------------
// GLM-compatible subset.
// ------------------------------------------------------------------------
struct dvec3
{
    double x, y, z;
    dvec3() : x(0), y(0), z(0) {}
    dvec3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit dvec3(double s) : x(s), y(s), z(s) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

static inline dvec3 operator+(const dvec3& a, const dvec3& b) { return dvec3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline dvec3 operator-(const dvec3& a, const dvec3& b) { return dvec3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline dvec3 operator-(const dvec3& a) { return dvec3(-a.x, -a.y, -a.z); }
static inline dvec3 operator*(const dvec3& a, double s) { return dvec3(a.x * s, a.y * s, a.z * s); }
static inline dvec3 operator*(double s, const dvec3& a) { return dvec3(s * a.x, s * a.y, s * a.z); }
static inline dvec3 operator/(const dvec3& a, double s) { return dvec3(a.x / s, a.y / s, a.z / s); }
static inline dvec3 operator/(const dvec3& a, const dvec3& b) { return dvec3(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline dvec3 operator*(const dvec3& a, const dvec3& b) { return dvec3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double dot(const dvec3& a, const dvec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double length(const dvec3& a) { return std::sqrt(dot(a, a)); }
static inline dvec3 normalize(const dvec3& a) { return a * (1.0 / std::sqrt(dot(a, a))); }
static inline dvec3 cross(const dvec3& x, const dvec3& y)
{
    return dvec3(x.y * y.z - y.y * x.z,
                 x.z * y.x - y.z * x.x,
                 x.x * y.y - y.x * x.y);
}
static inline double gmin(double a, double b) { return (b < a) ? b : a; }
static inline double gmax(double a, double b) { return (a < b) ? b : a; }
static inline double gsign(double x) { return (double)(0.0 < x) - (double)(x < 0.0); }

// Column-major 3x3: c[col][row].
struct dmat3
{
    dvec3 c[3];
    dmat3() {}
    explicit dmat3(double s) { c[0] = dvec3(s, 0, 0); c[1] = dvec3(0, s, 0); c[2] = dvec3(0, 0, s); }
    dmat3(const dvec3& a, const dvec3& b, const dvec3& d) { c[0] = a; c[1] = b; c[2] = d; }
    dvec3& operator[](int i) { return c[i]; }
    const dvec3& operator[](int i) const { return c[i]; }
};

static inline dmat3 operator*(const dmat3& A, const dmat3& B)
{
    dmat3 R;
    for(int j = 0; j < 3; ++j)
        for(int i = 0; i < 3; ++i)
            R[j][i] = A[0][i] * B[j][0] + A[1][i] * B[j][1] + A[2][i] * B[j][2];
    return R;
}
static inline dvec3 operator*(const dmat3& m, const dvec3& v)
{
    return dvec3(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z,
                 m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
                 m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}
static inline dmat3 operator+(const dmat3& A, const dmat3& B) { return dmat3(A[0] + B[0], A[1] + B[1], A[2] + B[2]); }
static inline dmat3 operator-(const dmat3& A, const dmat3& B) { return dmat3(A[0] - B[0], A[1] - B[1], A[2] - B[2]); }
static inline dmat3 operator*(const dmat3& A, double s) { return dmat3(A[0] * s, A[1] * s, A[2] * s); }
static inline dmat3 operator*(double s, const dmat3& A) { return dmat3(s * A[0], s * A[1], s * A[2]); }
static inline dmat3 transpose(const dmat3& m)
{
    dmat3 R;
    for(int j = 0; j < 3; ++j)
        for(int i = 0; i < 3; ++i)
            R[j][i] = m[i][j];
    return R;
}
static inline double determinant(const dmat3& m)
{
    return + m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
           - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
           + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]);
}
static inline dmat3 inverse(const dmat3& m)
{
    double oneOverDet = 1.0 / (
        + m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
        - m[1][0] * (m[0][1] * m[2][2] - m[2][1] * m[0][2])
        + m[2][0] * (m[0][1] * m[1][2] - m[1][1] * m[0][2]));
    dmat3 I;
    I[0][0] = + (m[1][1] * m[2][2] - m[2][1] * m[1][2]) * oneOverDet;
    I[1][0] = - (m[1][0] * m[2][2] - m[2][0] * m[1][2]) * oneOverDet;
    I[2][0] = + (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * oneOverDet;
    I[0][1] = - (m[0][1] * m[2][2] - m[2][1] * m[0][2]) * oneOverDet;
    I[1][1] = + (m[0][0] * m[2][2] - m[2][0] * m[0][2]) * oneOverDet;
    I[2][1] = - (m[0][0] * m[2][1] - m[2][0] * m[0][1]) * oneOverDet;
    I[0][2] = + (m[0][1] * m[1][2] - m[1][1] * m[0][2]) * oneOverDet;
    I[1][2] = - (m[0][0] * m[1][2] - m[1][0] * m[0][2]) * oneOverDet;
    I[2][2] = + (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * oneOverDet;
    return I;
}
// glm::mix for matrices, x + a (y - x) component-wise.
static inline dmat3 mixm(const dmat3& x, const dmat3& y, double a) { return x + a * (y - x); }

// ------------------------------------------------------------------------
// libm selection.
// ------------------------------------------------------------------------
static int g_libm = 1;
static inline double pow23(double x) { return g_libm ? gm_pow23(x) : std::pow(x, 2.0 / 3.0); }
static inline double pow13(double x) { return g_libm ? gm_cbrt_nonneg(x) : std::pow(x, 1.0 / 3.0); }
static inline double dexp(double x) { return g_libm ? gm_exp(x) : std::exp(x); }
static inline double dcos(double x) { return g_libm ? gm_cos(x) : std::cos(x); }
static inline double dpow(double x, double y) { return g_libm ? gm_pow(x, y) : std::pow(x, y); }

// ------------------------------------------------------------------------
// Mesh model (DataStructures/Mesh.h:25-267).
// ------------------------------------------------------------------------
enum { TET = 0, PYR = 1, PRI = 2, HEX = 3 };

// MeshTet::edges / MeshPri::edges / MeshHex::edges (Mesh.cpp:21-28, 70-80, 98-111).
static const int TET_EDGES[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
static const int PRI_EDGES[9][2] = {{0,1},{0,2},{1,2},{0,3},{1,4},{2,5},{3,4},{3,5},{4,5}};
static const int HEX_EDGES[12][2] = {{0,1},{0,3},{1,2},{2,3},{0,4},{1,5},{2,6},{3,7},{4,5},{4,7},{5,6},{6,7}};

struct NeigElem { uint32_t id; int type; int vId; };

// PlaneConstraint (FaceConstraint.cpp:173-185) and SegmentConstraint
// (EdgeConstraint.cpp:177-199) with their constructors' normalisations.
struct SlideConstraint
{
    int kind = 0;          // 1 plane, 2 segment
    dvec3 p, n;            // plane: point, unit normal
    dvec3 a, b, u;         // segment: ends, unit direction
    double length = 0.0;
};

struct Topo
{
    std::vector<uint32_t> neighborVerts;
    std::vector<NeigElem> neighborElems;
    int constraint;       // 0: free (VolumeConstraint), -1: fixed (VertexConstraint), k + 1: slides on constraint k
    dvec3 fixedPos;       // VertexConstraint::_pos (VertexConstraint.cpp:60-63)
    int fixedId = -1;     // identity of the VertexConstraint object (one per fixed vertex)
    const SlideConstraint* slide = nullptr;
};

struct Range { size_t begin = 0, end = 0; };

struct ParallelGroup
{
    Range boundaryRange, subsurfaceRange, interiorRange;
    std::vector<uint32_t> undispatchedNodes;
};

// Crew selections.
enum { EVAL_MEAN_RATIO = 0, EVAL_METRIC_CONFORMITY = 1 };
enum { MEAS_METRIC_FREE = 0, MEAS_METRIC_WISE = 1 };
enum { SAMP_UNIFORM = 0, SAMP_ANALYTIC_SHELL = 1, SAMP_TEXTURE = 2, SAMP_ANALYTIC = 3 };

struct Crew
{
    int evaluator = EVAL_MEAN_RATIO;
    int measurer = MEAS_METRIC_FREE;
    int sampler = SAMP_UNIFORM;
    double scaling = 1.0, scaling2 = 1.0, scaling3 = 1.0;   // AbstractSampler.cpp:49-54
    // Texture grid (TextureSampler.cpp:36-70): FP32 texels, 6 per cell
    // (m00 m01 m02 m11 m12 m22), cell (i,j,k) at index ((k*ny)+j)*nx+i.
    int tsize[3] = {0, 0, 0};
    dvec3 tmin, text;
    std::vector<float> texels;
    // AbstractSampler's aspect ratio and rotation (AbstractSampler.cpp:17-41).
    double aspect = 1.0;
    dmat3 rotMat, rotInv;
};

// glm::rotate (GLM's matrix_transform) restricted to the 3x3 block:
// m times the rotation by `angle` about the normalised axis.
static dmat3 glmRotate(const dmat3& m, double angle, dvec3 v)
{
    const double a = angle;
    const double c = std::cos(a);
    const double s = std::sin(a);
    dvec3 axis = v * (1.0 / std::sqrt(dot(v, v)));
    dvec3 temp = axis * (1.0 - c);
    dmat3 R;
    R[0][0] = c + temp[0] * axis[0];
    R[0][1] = temp[0] * axis[1] + s * axis[2];
    R[0][2] = temp[0] * axis[2] - s * axis[1];
    R[1][0] = temp[1] * axis[0] - s * axis[2];
    R[1][1] = c + temp[1] * axis[1];
    R[1][2] = temp[1] * axis[2] + s * axis[0];
    R[2][0] = temp[2] * axis[0] + s * axis[1];
    R[2][1] = temp[2] * axis[1] - s * axis[0];
    R[2][2] = c + temp[2] * axis[2];
    dmat3 out;
    out[0] = m[0] * R[0][0] + m[1] * R[0][1] + m[2] * R[0][2];
    out[1] = m[0] * R[1][0] + m[1] * R[1][1] + m[2] * R[1][2];
    out[2] = m[0] * R[2][0] + m[1] * R[2][1] + m[2] * R[2][2];
    return out;
}

// AbstractSampler's constructor (AbstractSampler.cpp:30-41).
static void samplerRotation(Crew& c)
{
    const double pi = 3.14159265358979323846;
    c.rotMat = dmat3(1.0);
    c.rotMat = glmRotate(c.rotMat, pi / 4.0, dvec3(0, 0, 1));
    c.rotMat = glmRotate(c.rotMat, pi / 4.0, dvec3(0, -1, 0));
    c.rotInv = transpose(c.rotMat);
}

struct Mesh
{
    std::vector<dvec3> verts;
    std::vector<Topo> topos;
    std::vector<std::array<uint32_t, 4>> tets;
    std::vector<std::array<uint32_t, 6>> pris;
    std::vector<std::array<uint32_t, 8>> hexs;

    // NodeGroups (NodeGroups.h:160-192)
    Range fixedNodes, boundaryNodes, subsurfaceNodes, interiorNodes;
    std::vector<uint32_t> serialGroup;
    std::vector<ParallelGroup> parallelGroups;
    std::vector<int64_t> permutation;   // positions[] of determinePositions: old id -> new id

    Crew crew;
    std::vector<SlideConstraint> slides;
};

// ------------------------------------------------------------------------
// Topology: Mesh::compileNeighborhoods / addEdge (Mesh.cpp:365-460).
// ------------------------------------------------------------------------
static void addEdge(Mesh& m, uint32_t a, uint32_t b)
{
    std::vector<uint32_t>& n = m.topos[a].neighborVerts;
    for(size_t i = 0; i < n.size(); ++i)
        if(n[i] == b)
            return;
    m.topos[a].neighborVerts.push_back(b);
    m.topos[b].neighborVerts.push_back(a);
}

static void compileNeighborhoods(Mesh& m)
{
    for(Topo& t : m.topos) { t.neighborElems.clear(); t.neighborVerts.clear(); }
    for(size_t i = 0; i < m.tets.size(); ++i)
    {
        for(int v = 0; v < 4; ++v)
            m.topos[m.tets[i][v]].neighborElems.push_back(NeigElem{(uint32_t)i, TET, v});
        for(int e = 0; e < 6; ++e)
            addEdge(m, m.tets[i][TET_EDGES[e][0]], m.tets[i][TET_EDGES[e][1]]);
    }
    for(size_t i = 0; i < m.pris.size(); ++i)
    {
        for(int v = 0; v < 6; ++v)
            m.topos[m.pris[i][v]].neighborElems.push_back(NeigElem{(uint32_t)i, PRI, v});
        for(int e = 0; e < 9; ++e)
            addEdge(m, m.pris[i][PRI_EDGES[e][0]], m.pris[i][PRI_EDGES[e][1]]);
    }
    for(size_t i = 0; i < m.hexs.size(); ++i)
    {
        for(int v = 0; v < 8; ++v)
            m.topos[m.hexs[i][v]].neighborElems.push_back(NeigElem{(uint32_t)i, HEX, v});
        for(int e = 0; e < 12; ++e)
            addEdge(m, m.hexs[i][HEX_EDGES[e][0]], m.hexs[i][HEX_EDGES[e][1]]);
    }
}

// ------------------------------------------------------------------------
// NodeGroups::build (NodeGroups.cpp:104-122) and its parts.
// ------------------------------------------------------------------------
static const int NO_GROUP = -2, UNSET_GROUP = -1;
static const int FIXED_TYPE = 0, BOUND_TYPE = 1, SUBSU_TYPE = 2, INTER_TYPE = 3;

static bool isFixed(const Topo& t) { return t.constraint < 0; }         // AbstractConstraint.h:65-68
static bool isConstrained(const Topo& t) { return t.constraint != 0; }  // AbstractConstraint.h:70-73
static bool isMovableBound(const Topo& t) { return isConstrained(t) && !isFixed(t); } // NodeGroups.cpp:695-704

// NodeGroups.cpp:124-241
static void determineTypes(const Mesh& m, std::vector<int>& types)
{
    size_t V = m.verts.size();
    types.assign(V, -1);
    std::vector<bool> bTets(m.tets.size(), false), bPris(m.pris.size(), false), bHexs(m.hexs.size(), false);
    for(size_t e = 0; e < m.tets.size(); ++e)
        for(int k = 0; k < 4; ++k)
            if(isMovableBound(m.topos[m.tets[e][k]])) { bTets[e] = true; break; }
    for(size_t e = 0; e < m.pris.size(); ++e)
        for(int k = 0; k < 6; ++k)
            if(isMovableBound(m.topos[m.pris[e][k]])) { bPris[e] = true; break; }
    for(size_t e = 0; e < m.hexs.size(); ++e)
        for(int k = 0; k < 8; ++k)
            if(isMovableBound(m.topos[m.hexs[e][k]])) { bHexs[e] = true; break; }
    for(size_t v = 0; v < V; ++v)
    {
        const Topo& t = m.topos[v];
        if(isFixed(t)) types[v] = FIXED_TYPE;
        else if(isConstrained(t)) types[v] = BOUND_TYPE;
        else
        {
            bool sub = false;
            for(const NeigElem& ne : t.neighborElems)
                if((ne.type == TET && bTets[ne.id]) || (ne.type == PRI && bPris[ne.id]) ||
                   (ne.type == HEX && bHexs[ne.id])) { sub = true; break; }
            types[v] = sub ? SUBSU_TYPE : INTER_TYPE;
        }
    }
}

// NodeGroups.cpp:243-363 (greedy BFS colouring with std::set, literally).
static void determineGroups(const Mesh& m, std::vector<int>& groups)
{
    size_t V = m.verts.size();
    size_t seekStart = 0;
    std::set<uint32_t> existingGroups;
    std::vector<size_t> nextNodes;
    groups.assign(V, NO_GROUP);
    while(nextNodes.size() < V)
    {
        size_t firstNode = nextNodes.size();
        for(size_t vId = seekStart; vId < V; ++vId)
        {
            ++seekStart;
            if(groups[vId] == NO_GROUP)
            {
                groups[vId] = UNSET_GROUP;
                nextNodes.push_back(vId);
                break;
            }
        }
        for(size_t v = firstNode; v < nextNodes.size(); ++v)
        {
            uint32_t vId = (uint32_t)nextNodes[v];
            const Topo& topo = m.topos[vId];
            std::set<uint32_t> availableGroups = existingGroups;
            for(size_t e = 0; e < topo.neighborElems.size(); ++e)
            {
                const NeigElem& ne = topo.neighborElems[e];
                const uint32_t* ev = ne.type == TET ? m.tets[ne.id].data()
                                   : ne.type == PRI ? m.pris[ne.id].data() : m.hexs[ne.id].data();
                int nv = ne.type == TET ? 4 : ne.type == PRI ? 6 : 8;
                for(int n = 0; n < nv; ++n)
                {
                    int& g = groups[ev[n]];
                    if(g == NO_GROUP) { g = UNSET_GROUP; nextNodes.push_back(ev[n]); }
                    else if(g != UNSET_GROUP) availableGroups.erase((uint32_t)g);
                }
            }
            int g;
            if(availableGroups.empty()) { g = (int)existingGroups.size(); existingGroups.insert((uint32_t)g); }
            else g = (int)*availableGroups.begin();
            groups[vId] = g;
        }
    }
}

// NodeGroups.cpp:365-450
static void determinePositions(Mesh& m, const std::vector<int>& types, const std::vector<int>& groups,
                               std::vector<int>& positions)
{
    size_t V = m.verts.size();
    m.fixedNodes = Range(); m.boundaryNodes = Range(); m.subsurfaceNodes = Range(); m.interiorNodes = Range();
    for(size_t v = 0; v < V; ++v)
    {
        if(types[v] == FIXED_TYPE) ++m.fixedNodes.end;
        else if(types[v] == BOUND_TYPE) ++m.boundaryNodes.end;
        else if(types[v] == SUBSU_TYPE) ++m.subsurfaceNodes.end;
        else if(types[v] == INTER_TYPE) ++m.interiorNodes.end;
    }
    m.boundaryNodes.begin += m.fixedNodes.end;   m.boundaryNodes.end += m.fixedNodes.end;
    m.subsurfaceNodes.begin += m.boundaryNodes.end; m.subsurfaceNodes.end += m.boundaryNodes.end;
    m.interiorNodes.begin += m.subsurfaceNodes.end; m.interiorNodes.end += m.subsurfaceNodes.end;

    std::vector<int> indices(V);
    std::iota(indices.begin(), indices.end(), 0);
    const std::vector<Topo>& topos = m.topos;
    std::sort(indices.begin(), indices.end(), [&](int a, int b) {
        if(types[a] == types[b])
        {
            if(types[a] == FIXED_TYPE) return false;
            if(groups[a] == groups[b])
                return topos[a].neighborElems.size() < topos[b].neighborElems.size();
            return groups[a] < groups[b];
        }
        return types[a] < types[b];
    });
    positions.resize(V);
    for(size_t v = 0; v < V; ++v)
        positions[indices[v]] = (int)v;
}

// NodeGroups.cpp:452-619 (the cycle-following in-place permutation).
static void clusterNodes(Mesh& m, std::vector<int>& types, std::vector<int>& groups, std::vector<int>& positions)
{
    size_t V = m.verts.size();
    for(auto& t : m.tets) for(int k = 0; k < 4; ++k) t[k] = (uint32_t)positions[t[k]];
    for(auto& p : m.pris) for(int k = 0; k < 6; ++k) p[k] = (uint32_t)positions[p[k]];
    for(auto& h : m.hexs) for(int k = 0; k < 8; ++k) h[k] = (uint32_t)positions[h[k]];
    for(Topo& t : m.topos) for(uint32_t& n : t.neighborVerts) n = (uint32_t)positions[n];

    m.serialGroup.clear();
    m.parallelGroups.clear();
    int lastType = -1, lastGroup = -1;
    for(size_t vId = 0; vId < V; ++vId)
    {
        while((size_t)positions[vId] != vId)
        {
            size_t nId = (size_t)positions[vId];
            std::swap(m.verts[vId], m.verts[nId]);
            std::swap(m.topos[vId], m.topos[nId]);
            std::swap(types[vId], types[nId]);
            std::swap(groups[vId], groups[nId]);
            std::swap(positions[vId], positions[nId]);
        }
        int type = types[vId];
        if(type != lastType) { lastType = type; lastGroup = -1; }
        if(type > FIXED_TYPE)
        {
            m.serialGroup.push_back((uint32_t)vId);
            int gId = groups[vId];
            if(gId != lastGroup)
            {
                lastGroup = gId;
                if((int)m.parallelGroups.size() <= gId) m.parallelGroups.resize(gId + 1);
                ParallelGroup& cg = m.parallelGroups[gId];
                if(type == BOUND_TYPE) { cg.boundaryRange.begin = vId; cg.boundaryRange.end = vId; }
                else if(type == SUBSU_TYPE) { cg.subsurfaceRange.begin = vId; cg.subsurfaceRange.end = vId; }
                else if(type == INTER_TYPE) { cg.interiorRange.begin = vId; cg.interiorRange.end = vId; }
            }
            ParallelGroup& g = m.parallelGroups[gId];
            g.undispatchedNodes.push_back((uint32_t)vId);
            if(type == BOUND_TYPE) ++g.boundaryRange.end;
            else if(type == SUBSU_TYPE) ++g.subsurfaceRange.end;
            else if(type == INTER_TYPE) ++g.interiorRange.end;
        }
    }
}

static void buildNodeGroups(Mesh& m)
{
    std::vector<int> types, groups, positions;
    determineTypes(m, types);
    determineGroups(m, groups);
    determinePositions(m, types, groups, positions);
    m.permutation.assign(positions.begin(), positions.end());
    clusterNodes(m, types, groups, positions);
}

// ------------------------------------------------------------------------
// Samplers.
// ------------------------------------------------------------------------
// Spherical-shell anisotropy, centre (0.5,0.5,0.5):
//   a = 1 + 9 exp(-50 (r - 0.3)^2),  M = K^2 (a^2 I + (1 - a^2) d d^T / r^2).
static dmat3 shellMetric(const Crew& c, const dvec3& p)
{
    double dx = p.x - 0.5, dy = p.y - 0.5, dz = p.z - 0.5;
    double r2 = dx * dx + dy * dy + dz * dz;
    double r = std::sqrt(r2);
    double u = r - 0.3;
    double a = 1.0 + 9.0 * dexp(-50.0 * (u * u));
    double a2 = a * a;
    double w = (r2 > 0.0) ? (1.0 - a2) / r2 : 0.0;
    double k2 = c.scaling2;
    double m00 = k2 * (a2 + w * (dx * dx));
    double m11 = k2 * (a2 + w * (dy * dy));
    double m22 = k2 * (a2 + w * (dz * dz));
    double m01 = k2 * (w * (dx * dy));
    double m02 = k2 * (w * (dx * dz));
    double m12 = k2 * (w * (dy * dz));
    return dmat3(dvec3(m00, m01, m02), dvec3(m01, m11, m12), dvec3(m02, m12, m22));
}

// AbstractSampler::vertMetric (AbstractSampler.cpp:190-208), the field of
// AnalyticSampler::metricAt (AnalyticSampler.cpp:45-49).
static dmat3 analyticMetric(const Crew& c, const dvec3& position)
{
    double x = (c.rotMat * position).x * (2.5 * 3.14159265358979323846);
    double a = c.aspect;
    double cc = (1.0 - dcos(x)) / 2.0;
    double sizeX = c.scaling * dpow(a, dpow(cc, a));
    double Mx = sizeX * sizeX;
    double My = c.scaling2;
    double Mz = My;
    dmat3 M(dvec3(Mx, 0, 0), dvec3(0, My, 0), dvec3(0, 0, Mz));
    return c.rotInv * M * c.rotMat;
}

static dmat3 texel(const Crew& c, int i, int j, int k)
{
    const float* t = &c.texels[6 * (((size_t)k * c.tsize[1] + j) * c.tsize[0] + i)];
    double m00 = t[0], m01 = t[1], m02 = t[2], m11 = t[3], m12 = t[4], m22 = t[5];
    return dmat3(dvec3(m00, m01, m02), dvec3(m01, m11, m12), dvec3(m02, m12, m22));
}

static int clampi(int v, int lo, int hi) { return std::min(std::max(v, lo), hi); }

// TextureSampler::metricAt (TextureSampler.cpp:450-488) and cellId (:513-524).
static dmat3 textureMetric(const Crew& c, const dvec3& position)
{
    dvec3 size((double)c.tsize[0], (double)c.tsize[1], (double)c.tsize[2]);
    dvec3 cs = c.text / size;
    dvec3 minB = c.tmin;
    dvec3 maxB = (c.tmin + c.text) - (cs * 1.5);
    dvec3 q = position + dvec3(-cs.x, -cs.y, -cs.z) / 2.0;
    dvec3 cp0(gmin(gmax(q.x, minB.x), maxB.x), gmin(gmax(q.y, minB.y), maxB.y), gmin(gmax(q.z, minB.z), maxB.z));
    dvec3 ratio = (cp0 - c.tmin) / c.text;
    dvec3 fid = ratio * size;
    int id0[3], id1[3];
    for(int d = 0; d < 3; ++d)
    {
        id0[d] = clampi((int)fid[d], 0, c.tsize[d] - 1);
        id1[d] = std::min(id0[d] + 1, c.tsize[d] - 1);
    }
    dmat3 m000 = texel(c, id0[0], id0[1], id0[2]);
    dmat3 m100 = texel(c, id1[0], id0[1], id0[2]);
    dmat3 m010 = texel(c, id0[0], id1[1], id0[2]);
    dmat3 m110 = texel(c, id1[0], id1[1], id0[2]);
    dmat3 m001 = texel(c, id0[0], id0[1], id1[2]);
    dmat3 m101 = texel(c, id1[0], id0[1], id1[2]);
    dmat3 m011 = texel(c, id0[0], id1[1], id1[2]);
    dmat3 m111 = texel(c, id1[0], id1[1], id1[2]);
    dvec3 c0Center = cs * (dvec3((double)id0[0], (double)id0[1], (double)id0[2]) + dvec3(0.5));
    dvec3 a = (position - (c.tmin + c0Center)) / cs;
    a = dvec3(gmin(gmax(a.x, 0.0), 1.0), gmin(gmax(a.y, 0.0), 1.0), gmin(gmax(a.z, 0.0), 1.0));
    dmat3 mx00 = mixm(m000, m100, a.x);
    dmat3 mx10 = mixm(m010, m110, a.x);
    dmat3 mx01 = mixm(m001, m101, a.x);
    dmat3 mx11 = mixm(m011, m111, a.x);
    dmat3 mxy0 = mixm(mx00, mx10, a.y);
    dmat3 mxy1 = mixm(mx01, mx11, a.y);
    return mixm(mxy0, mxy1, a.z);
}

// AbstractSampler::metricAt dispatch: UniformSampler.cpp:45-51 (K^2 I),
// the analytic field, TextureSampler.cpp:450-488.
static dmat3 metricAt(const Crew& c, const dvec3& p)
{
    switch(c.sampler)
    {
    case SAMP_ANALYTIC_SHELL: return shellMetric(c, p);
    case SAMP_ANALYTIC: return analyticMetric(c, p);
    case SAMP_TEXTURE: return textureMetric(c, p);
    default: return dmat3(c.scaling * c.scaling);
    }
}

// ------------------------------------------------------------------------
// Measurers.
// ------------------------------------------------------------------------
static const double DIFF_THRESHOLD = 0.01;   // MetricWiseMeasurer.cpp:9

// MetricWiseMeasurer::riemannianDistance, localized division (:65-116).
static double riemannianDistanceWise(const Crew& c, const dvec3& a, const dvec3& b)
{
    int curr = 0, base = 2;
    double len = 0.0;
    dvec3 d = b - a;
    dvec3 bv = d / (double)base;
    while(curr < base)
    {
        double p0 = (curr + 0.5) / base;
        double p1 = (curr + 1.5) / base;
        dmat3 M0 = metricAt(c, a + p0 * d);
        dmat3 M1 = metricAt(c, a + p1 * d);
        double l0 = std::sqrt(dot(bv, M0 * bv));
        double l1 = std::sqrt(dot(bv, M1 * bv));
        double sum = (l0 + l1);
        double diff = std::abs(l0 - l1) / (sum / 2.0);
        // Not in the reference: a NaN difference (zero-length pair or
        // non-finite input) would subdivide until the int indices overflow;
        // the product ends the walk here, and so does the oracle.
        if(!(diff == diff))
            return len + sum;
        if(diff < DIFF_THRESHOLD)
        {
            len += sum;
            curr += 2;
            if((curr & 2) == 0) { base >>= 1; curr >>= 1; }
            bv = d / (double)base;
        }
        else
        {
            base <<= 1;
            curr <<= 1;
            bv = bv / 2.0;
        }
    }
    return len;
}

// MetricFreeMeasurer::riemannianSegment (:36-43) / MetricWise (:118-126).
static dvec3 riemannianSegment(const Crew& c, const dvec3& a, const dvec3& b)
{
    if(c.measurer == MEAS_METRIC_WISE)
        return normalize(b - a) * riemannianDistanceWise(c, a, b);
    return (b - a) * c.scaling;
}

// MetricFreeMeasurer::tetVolume (:45-56) / MetricWise (:128-139).
static double tetVolume(const Crew& c, const dvec3 vp[])
{
    if(c.measurer == MEAS_METRIC_WISE)
    {
        double det = determinant(dmat3(riemannianSegment(c, vp[0], vp[3]),
                                       riemannianSegment(c, vp[1], vp[3]),
                                       riemannianSegment(c, vp[2], vp[3])));
        return det / 6.0;
    }
    double det = determinant(dmat3(vp[3] - vp[0], vp[3] - vp[1], vp[3] - vp[2]));
    return c.scaling3 * det / 6.0;
}

// MetricFreeMeasurer::priVolume (:58-75) / MetricWise (:141-158).
static double priVolume(const Crew& c, const dvec3 vp[])
{
    dvec3 e02, e12, e32, e42, e52;
    if(c.measurer == MEAS_METRIC_WISE)
    {
        e02 = riemannianSegment(c, vp[0], vp[2]);
        e12 = riemannianSegment(c, vp[1], vp[2]);
        e32 = riemannianSegment(c, vp[3], vp[2]);
        e42 = riemannianSegment(c, vp[4], vp[2]);
        e52 = riemannianSegment(c, vp[5], vp[2]);
    }
    else
    {
        e02 = vp[2] - vp[0];
        e12 = vp[2] - vp[1];
        e32 = vp[2] - vp[3];
        e42 = vp[2] - vp[4];
        e52 = vp[2] - vp[5];
    }
    double detSum = 0.0;
    detSum += determinant(dmat3(e32, e52, e42));
    detSum += determinant(dmat3(e02, e32, e42));
    detSum += determinant(dmat3(e12, e02, e42));
    return c.measurer == MEAS_METRIC_WISE ? detSum / 6.0 : c.scaling3 * detSum / 6.0;
}

// MetricFreeMeasurer::hexVolume (:77-105) / MetricWise (:160-188).
static double hexVolume(const Crew& c, const dvec3 vp[])
{
    double detSum = 0.0;
    if(c.measurer == MEAS_METRIC_WISE)
    {
        detSum += determinant(dmat3(riemannianSegment(c, vp[1], vp[0]), riemannianSegment(c, vp[4], vp[0]), riemannianSegment(c, vp[3], vp[0])));
        detSum += determinant(dmat3(riemannianSegment(c, vp[1], vp[2]), riemannianSegment(c, vp[3], vp[2]), riemannianSegment(c, vp[6], vp[2])));
        detSum += determinant(dmat3(riemannianSegment(c, vp[4], vp[5]), riemannianSegment(c, vp[1], vp[5]), riemannianSegment(c, vp[6], vp[5])));
        detSum += determinant(dmat3(riemannianSegment(c, vp[4], vp[7]), riemannianSegment(c, vp[6], vp[7]), riemannianSegment(c, vp[3], vp[7])));
        detSum += determinant(dmat3(riemannianSegment(c, vp[1], vp[4]), riemannianSegment(c, vp[6], vp[4]), riemannianSegment(c, vp[3], vp[4])));
        return detSum / 6.0;
    }
    detSum += determinant(dmat3(vp[0] - vp[1], vp[0] - vp[4], vp[0] - vp[3]));
    detSum += determinant(dmat3(vp[2] - vp[1], vp[2] - vp[3], vp[2] - vp[6]));
    detSum += determinant(dmat3(vp[5] - vp[4], vp[5] - vp[1], vp[5] - vp[6]));
    detSum += determinant(dmat3(vp[7] - vp[4], vp[7] - vp[6], vp[7] - vp[3]));
    detSum += determinant(dmat3(vp[4] - vp[1], vp[4] - vp[6], vp[4] - vp[3]));
    return c.scaling3 * detSum / 6.0;
}

// AbstractMeasurer::computeLocalElementSize (AbstractMeasurer.cpp:137-155).
static double computeLocalElementSize(const Mesh& m, uint32_t vId)
{
    const dvec3& pos = m.verts[vId];
    const std::vector<uint32_t>& nv = m.topos[vId].neighborVerts;
    double total = 0.0;
    for(size_t n = 0; n < nv.size(); ++n)
        total += length(pos - m.verts[nv[n]]);
    return total / (double)nv.size();
}

// ------------------------------------------------------------------------
// Evaluators.
// ------------------------------------------------------------------------
// AbstractEvaluator.cpp:28-31
static const dmat3 Fr_TET_INV(
    dvec3(1, 0, 0),
    dvec3(-0.5773502691896257645091, 1.154700538379251529018, 0),
    dvec3(-0.4082482904638630163662, -0.4082482904638630163662, 1.224744871391589049099));
// AbstractEvaluator.cpp:32-35
static const dmat3 Fr_PRI_INV(
    dvec3(1.0, 0.0, 0.0),
    dvec3(-0.5773502691896257645091, 1.154700538379251529018, 0.0),
    dvec3(0.0, 0.0, 1.0));

// MeanRatioEvaluator::cornerQuality (MeanRatioEvaluator.cpp:22-32).
static double cornerQuality(const dmat3& Fk)
{
    double det = determinant(Fk);
    double frob2 = dot(Fk[0], Fk[0]) + dot(Fk[1], Fk[1]) + dot(Fk[2], Fk[2]);
    return gsign(det) * 3.0 * pow23(std::abs(det)) / frob2;
}

// MetricConformityEvaluator::specifiedMetric (:25-45).
static dmat3 specifiedMetric(const Crew& c, const dvec3& v0, const dvec3& v1, const dvec3& v2, const dvec3& v3)
{
    const double H = 0.5;
    const double Q = (1.0 - H) / 3.0;
    return metricAt(c, (v0 + v1 + v2 + v3) / 4.0) * (-0.8) +
           metricAt(c, v0 * H + v1 * Q + v2 * Q + v3 * Q) * 0.45 +
           metricAt(c, v0 * Q + v1 * H + v2 * Q + v3 * Q) * 0.45 +
           metricAt(c, v0 * Q + v1 * Q + v2 * H + v3 * Q) * 0.45 +
           metricAt(c, v0 * Q + v1 * Q + v2 * Q + v3 * H) * 0.45;
}

// MetricConformityEvaluator::metricConformity (:47-72).
static double metricConformity(const dmat3& Fk, const dmat3& Ms)
{
    dmat3 Fk_inv = inverse(transpose(Fk));
    dmat3 Mk = Fk_inv * transpose(Fk_inv);
    dmat3 Mk_inv = inverse(Mk);
    dmat3 Ms_inv = inverse(Ms);
    dmat3 tNc = Mk_inv * Ms + Ms_inv * Mk - dmat3(2.0);
    double f2 = dot(tNc[0], tNc[0]) + dot(tNc[1], tNc[1]) + dot(tNc[2], tNc[2]);
    if(!std::isnan(f2))
        return gsign(determinant(Fk)) / std::sqrt(1.0 + std::sqrt(f2));
    return 0.0;
}

// MeanRatioEvaluator::tetQuality (:34-50), MetricConformityEvaluator::tetQuality (:74-91).
static double tetQuality(const Crew& c, const dvec3 vp[])
{
    if(c.evaluator == EVAL_METRIC_CONFORMITY)
    {
        dvec3 e03 = vp[3] - vp[0], e13 = vp[3] - vp[1], e23 = vp[3] - vp[2];
        dmat3 Fk = dmat3(e03, e13, e23) * Fr_TET_INV;
        dmat3 Ms0 = specifiedMetric(c, vp[0], vp[1], vp[2], vp[3]);
        return metricConformity(Fk, Ms0);
    }
    dvec3 e03 = riemannianSegment(c, vp[0], vp[3]);
    dvec3 e13 = riemannianSegment(c, vp[1], vp[3]);
    dvec3 e23 = riemannianSegment(c, vp[2], vp[3]);
    dmat3 Fk0 = dmat3(e03, e13, e23) * Fr_TET_INV;
    return cornerQuality(Fk0);
}

static double hexMeanOrMin(const double q[8])
{
    // MeanRatioEvaluator.cpp:146-166 / MetricConformityEvaluator.cpp:203-223
    if(q[0] <= 0.0 || q[1] <= 0.0 || q[2] <= 0.0 || q[3] <= 0.0 ||
       q[4] <= 0.0 || q[5] <= 0.0 || q[6] <= 0.0 || q[7] <= 0.0)
    {
        double mn = q[0];
        for(int k = 1; k < 8; ++k) if(q[k] < mn) mn = q[k];
        return mn;
    }
    return 8.0 / (1 / q[0] + 1 / q[1] + 1 / q[2] + 1 / q[3] + 1 / q[4] + 1 / q[5] + 1 / q[6] + 1 / q[7]);
}

// MeanRatioEvaluator.cpp:85-103 / MetricConformityEvaluator.cpp:133-151
static double priMeanOrMin(const double q[6])
{
    if(q[0] <= 0.0 || q[1] <= 0.0 || q[2] <= 0.0 || q[3] <= 0.0 || q[4] <= 0.0 || q[5] <= 0.0)
    {
        double mn = q[0];
        for(int k = 1; k < 6; ++k) if(q[k] < mn) mn = q[k];
        return mn;
    }
    return 6.0 / (1 / q[0] + 1 / q[1] + 1 / q[2] + 1 / q[3] + 1 / q[4] + 1 / q[5]);
}

// MeanRatioEvaluator::priQuality (:52-104), MetricConformityEvaluator::priQuality (:93-152).
static double priQuality(const Crew& c, const dvec3 vp[])
{
    double q[6];
    dvec3 e03, e14, e25, e01, e12, e20, e34, e45, e53;
    if(c.evaluator == EVAL_METRIC_CONFORMITY)
    {
        e03 = vp[3] - vp[0]; e14 = vp[4] - vp[1]; e25 = vp[5] - vp[2];
        e01 = vp[1] - vp[0]; e12 = vp[2] - vp[1]; e20 = vp[0] - vp[2];
        e34 = vp[4] - vp[3]; e45 = vp[5] - vp[4]; e53 = vp[3] - vp[5];
    }
    else
    {
        e03 = riemannianSegment(c, vp[0], vp[3]);
        e14 = riemannianSegment(c, vp[1], vp[4]);
        e25 = riemannianSegment(c, vp[2], vp[5]);
        e01 = riemannianSegment(c, vp[0], vp[1]);
        e12 = riemannianSegment(c, vp[1], vp[2]);
        e20 = riemannianSegment(c, vp[2], vp[0]);
        e34 = riemannianSegment(c, vp[3], vp[4]);
        e45 = riemannianSegment(c, vp[4], vp[5]);
        e53 = riemannianSegment(c, vp[5], vp[3]);
    }
    dmat3 Fk[6] = {dmat3(-e01, e20, e03) * Fr_PRI_INV, dmat3(-e12, e01, e14) * Fr_PRI_INV,
                   dmat3(-e20, e12, e25) * Fr_PRI_INV, dmat3(-e34, e53, e03) * Fr_PRI_INV,
                   dmat3(-e45, e34, e14) * Fr_PRI_INV, dmat3(-e53, e45, e25) * Fr_PRI_INV};
    if(c.evaluator == EVAL_METRIC_CONFORMITY)
    {
        static const int T[6][4] = {{0,1,2,3},{0,1,2,4},{0,1,2,5},{0,3,4,5},{1,3,4,5},{2,3,4,5}};
        for(int k = 0; k < 6; ++k)
            q[k] = metricConformity(Fk[k], specifiedMetric(c, vp[T[k][0]], vp[T[k][1]], vp[T[k][2]], vp[T[k][3]]));
    }
    else
        for(int k = 0; k < 6; ++k) q[k] = cornerQuality(Fk[k]);
    return priMeanOrMin(q);
}

// MeanRatioEvaluator::hexQuality (:106-167), MetricConformityEvaluator::hexQuality (:154-224).
static double hexQuality(const Crew& c, const dvec3 vp[])
{
    double q[8];
    if(c.evaluator == EVAL_METRIC_CONFORMITY)
    {
        dvec3 e01 = vp[1] - vp[0], e03 = vp[3] - vp[0], e04 = vp[4] - vp[0], e12 = vp[2] - vp[1];
        dvec3 e15 = vp[5] - vp[1], e23 = vp[3] - vp[2], e26 = vp[6] - vp[2], e37 = vp[7] - vp[3];
        dvec3 e45 = vp[5] - vp[4], e47 = vp[7] - vp[4], e56 = vp[6] - vp[5], e67 = vp[7] - vp[6];
        dmat3 Fk[8] = {dmat3(e01, e04, -e03), dmat3(e01, e12, e15), dmat3(e12, e26, -e23), dmat3(e03, e23, e37),
                       dmat3(e04, e45, e47), dmat3(e15, -e56, e45), dmat3(e26, e56, e67), dmat3(e37, e67, -e47)};
        static const int T[8][4] = {{0,1,3,4},{0,1,2,5},{1,2,3,6},{0,2,3,7},{0,4,5,7},{1,4,5,6},{2,5,6,7},{3,4,6,7}};
        for(int k = 0; k < 8; ++k)
            q[k] = metricConformity(Fk[k], specifiedMetric(c, vp[T[k][0]], vp[T[k][1]], vp[T[k][2]], vp[T[k][3]]));
        return hexMeanOrMin(q);
    }
    dvec3 e01 = riemannianSegment(c, vp[0], vp[1]);
    dvec3 e03 = riemannianSegment(c, vp[0], vp[3]);
    dvec3 e04 = riemannianSegment(c, vp[0], vp[4]);
    dvec3 e12 = riemannianSegment(c, vp[1], vp[2]);
    dvec3 e15 = riemannianSegment(c, vp[1], vp[5]);
    dvec3 e23 = riemannianSegment(c, vp[2], vp[3]);
    dvec3 e26 = riemannianSegment(c, vp[2], vp[6]);
    dvec3 e37 = riemannianSegment(c, vp[3], vp[7]);
    dvec3 e45 = riemannianSegment(c, vp[4], vp[5]);
    dvec3 e47 = riemannianSegment(c, vp[4], vp[7]);
    dvec3 e56 = riemannianSegment(c, vp[5], vp[6]);
    dvec3 e67 = riemannianSegment(c, vp[6], vp[7]);
    dmat3 Fk[8] = {dmat3(e01, e04, -e03), dmat3(e01, e12, e15), dmat3(e12, e26, -e23), dmat3(e03, e23, e37),
                   dmat3(e04, e45, e47), dmat3(e15, -e56, e45), dmat3(e26, e56, e67), dmat3(e37, e67, -e47)};
    for(int k = 0; k < 8; ++k)
        q[k] = cornerQuality(Fk[k]);
    return hexMeanOrMin(q);
}

static double tetQualityMesh(const Mesh& m, uint32_t e)
{
    dvec3 vp[4];
    for(int k = 0; k < 4; ++k) vp[k] = m.verts[m.tets[e][k]];
    return tetQuality(m.crew, vp);
}

static double priQualityMesh(const Mesh& m, uint32_t e)
{
    dvec3 vp[6];
    for(int k = 0; k < 6; ++k) vp[k] = m.verts[m.pris[e][k]];
    return priQuality(m.crew, vp);
}

static double hexQualityMesh(const Mesh& m, uint32_t e)
{
    dvec3 vp[8];
    for(int k = 0; k < 8; ++k) vp[k] = m.verts[m.hexs[e][k]];
    return hexQuality(m.crew, vp);
}

// AbstractEvaluator::accumulatePatchQuality / finalizePatchQuality (:778-803).
static inline void accumulatePatchQuality(double& pq, double& pw, double q)
{
    if(pw != -1.0 && q > 0.0) { pw += 1.0; pq += 1 / q; }
    else { pw = -1.0; pq = std::min(pq, q); }
}
static inline double finalizePatchQuality(double pq, double pw)
{
    return (pw > 0.0) ? pw / pq : pq;
}

// Test hook: while set, every patch evaluation appends (x, y, z, quality)
// of the evaluated vertex, so tests can follow an optimiser's trial points.
static std::vector<double>* g_trace = nullptr;

static double patchQualityUntraced(const Mesh& m, uint32_t vId);

// AbstractEvaluator::patchQuality (:236-279).
static double patchQuality(const Mesh& m, uint32_t vId, int64_t* counter)
{
    if(counter) ++*counter;
    double q = patchQualityUntraced(m, vId);
    if(g_trace)
    {
        const dvec3& p = m.verts[vId];
        g_trace->insert(g_trace->end(), {p.x, p.y, p.z, q});
    }
    return q;
}

static double patchQualityUntraced(const Mesh& m, uint32_t vId)
{
    const Topo& topo = m.topos[vId];
    double pw = 0.0, pq = 0.0;
    for(const NeigElem& ne : topo.neighborElems)
    {
        if(ne.type == TET) accumulatePatchQuality(pq, pw, tetQualityMesh(m, ne.id));
        else if(ne.type == PRI) accumulatePatchQuality(pq, pw, priQualityMesh(m, ne.id));
        else if(ne.type == HEX) accumulatePatchQuality(pq, pw, hexQualityMesh(m, ne.id));
    }
    return finalizePatchQuality(pq, pw);
}

// ------------------------------------------------------------------------
// QualityHistogram (QualityHistogram.cpp:63-92).
// ------------------------------------------------------------------------
struct Histogram
{
    int32_t buckets[20] = {0};
    int64_t n = 0;
    double minQ = 1.0;
    double invSum = 0.0;

    // size_t(value * 20) clamped to [0, 19].  Values whose product truncates
    // to a negative integer (and NaN) land in the last bucket: that is what
    // the x86-64 double->size_t conversion produces before the clamp.
    static int bucketOf(double value)
    {
        double v = value * 20.0;
        if(v != v || v <= -1.0 || v >= 9.2e18) return 19;
        int64_t t = (int64_t)v;
        return (int)std::min<int64_t>(std::max<int64_t>(t, 0), 19);
    }
    void add(double value)
    {
        ++n;
        minQ = gmin(minQ, value);
        if(value > 0.0) invSum += 1.0 / value;
        else invSum = INFINITY;
        buckets[bucketOf(value)] += 1;
    }
    void merge(const Histogram& h)
    {
        if(h.n == 0) return;
        n += h.n;
        minQ = gmin(minQ, h.minQ);
        invSum += h.invSum;
        for(int i = 0; i < 20; ++i) buckets[i] += h.buckets[i];
    }
    double harmonicMean() const { return (double)n / invSum; }
};

// evaluateMeshQualitySerial (AbstractEvaluator.cpp:384-402) when chunks == 1;
// evaluateMeshQualityThread's chunking (:404-453) for chunks > 1.
static Histogram evaluateMesh(const Mesh& m, int chunks)
{
    Histogram out;
    int tetCount = (int)m.tets.size(), priCount = (int)m.pris.size(), hexCount = (int)m.hexs.size();
    if(chunks <= 1)
    {
        for(int i = 0; i < tetCount; ++i) out.add(tetQualityMesh(m, i));
        for(int i = 0; i < priCount; ++i) out.add(priQualityMesh(m, i));
        for(int i = 0; i < hexCount; ++i) out.add(hexQualityMesh(m, i));
        return out;
    }
    if(tetCount + priCount + hexCount == 0) return out;
    std::vector<Histogram> parts(chunks);
    std::vector<std::thread> th;
    for(int t = 0; t < chunks; ++t)
        th.emplace_back([&, t]() {
            int tb = (int)(((int64_t)tetCount * t) / chunks), te = (int)(((int64_t)tetCount * (t + 1)) / chunks);
            int pb = (int)(((int64_t)priCount * t) / chunks), pe = (int)(((int64_t)priCount * (t + 1)) / chunks);
            int hb = (int)(((int64_t)hexCount * t) / chunks), he = (int)(((int64_t)hexCount * (t + 1)) / chunks);
            for(int i = tb; i < te; ++i) parts[t].add(tetQualityMesh(m, i));
            for(int i = pb; i < pe; ++i) parts[t].add(priQualityMesh(m, i));
            for(int i = hb; i < he; ++i) parts[t].add(hexQualityMesh(m, i));
        });
    for(auto& t : th) t.join();
    for(int t = 0; t < chunks; ++t) out.merge(parts[t]);
    return out;
}

// ------------------------------------------------------------------------
// Constraint snapping (VertexConstraint.cpp:60-63; VolumeConstraint is the identity).
// ------------------------------------------------------------------------
static inline dvec3 snap(const Topo& t, const dvec3& pos)
{
    if(isFixed(t)) return t.fixedPos;
    const SlideConstraint* c = t.slide;
    if(!c) return pos;
    if(c->kind == 1)
    {
        // PlaneConstraint::operator() (FaceConstraint.cpp:180-185)
        dvec3 dist = pos - c->p;
        double l = dot(dist, c->n);
        return pos - c->n * l;
    }
    // SegmentConstraint::operator() (EdgeConstraint.cpp:185-199)
    dvec3 dist = pos - c->a;
    double l = dot(dist, c->u);
    if(l < 0) return c->a;
    else if(l > c->length) return c->b;
    dvec3 t2 = dist - c->u * l;
    return pos - t2;
}

// ------------------------------------------------------------------------
// Vertex-wise smoothers.
// ------------------------------------------------------------------------
struct SmoothParams
{
    int smoother = 0;            // 0 NM, 1 GD, 2 spawn search, 3 GETMe
    double spawnMoveCoeff = 1.0 / 20.0;
    std::vector<int64_t>* evalCount = nullptr;   // patch evaluations per vertex
};

struct dvec4 { dvec3 p; double w; };

// NelderMeadSmoother::smoothVertices (NelderMeadSmoother.cpp:100-261), constants :17-23.
static void nelderMeadVertex(Mesh& m, uint32_t vId, int64_t* cnt)
{
    const double NMValueConvergence = 0.000100;
    const int NMSecurityCycleCount = 12;
    const double NMLocalSizeToNodeShift = 1.0 / 24.0;
    const double NMAlpha = 1.0, NMBeta = 0.5, NMGamma = 2.0;

    double localSize = computeLocalElementSize(m, vId);
    double nodeShift = localSize * NMLocalSizeToNodeShift;
    dvec3& pos = m.verts[vId];
    const Topo& topo = m.topos[vId];
    bool constrained = isConstrained(topo);
    dvec4 vo{pos, patchQuality(m, vId, cnt)};
    dvec4 simplex[4] = {
        {pos + dvec3(nodeShift, 0, 0), 0},
        {pos + dvec3(0, nodeShift, 0), 0},
        {pos + dvec3(0, 0, nodeShift), 0},
        vo};
    int cycle = 0;
    bool reset = false, terminated = false;
    while(!terminated)
    {
        for(int p = 0; p < 3; ++p)
        {
            pos = simplex[p].p;
            if(constrained) pos = snap(topo, pos);
            simplex[p] = dvec4{pos, patchQuality(m, vId, cnt)};
        }
        if(simplex[0].w > simplex[1].w) std::swap(simplex[0], simplex[1]);
        if(simplex[1].w > simplex[2].w) std::swap(simplex[1], simplex[2]);
        if(simplex[2].w > simplex[3].w) std::swap(simplex[2], simplex[3]);
        if(simplex[0].w > simplex[1].w) std::swap(simplex[0], simplex[1]);
        if(simplex[1].w > simplex[2].w) std::swap(simplex[1], simplex[2]);
        if(simplex[0].w > simplex[1].w) std::swap(simplex[0], simplex[1]);

        for(; cycle < NMSecurityCycleCount; ++cycle)
        {
            dvec3 c = 1 / 3.0 * (simplex[1].p + simplex[2].p + simplex[3].p);
            double f = 0.0;
            pos = c + NMAlpha * (c - simplex[0].p);
            if(constrained) pos = snap(topo, pos);
            double fr = f = patchQuality(m, vId, cnt);
            dvec3 xr = pos;
            if(simplex[3].w < fr)
            {
                pos = c + NMGamma * (pos - c);
                if(constrained) pos = snap(topo, pos);
                double fe = f = patchQuality(m, vId, cnt);
                if(fe <= fr) { pos = xr; f = fr; }
            }
            else if(simplex[1].w >= fr)
            {
                if(fr > simplex[0].w)
                {
                    pos = c + NMBeta * (xr - c);
                    if(constrained) pos = snap(topo, pos);
                    f = patchQuality(m, vId, cnt);
                }
                else
                {
                    pos = c + NMBeta * (simplex[0].p - c);
                    if(constrained) pos = snap(topo, pos);
                    f = patchQuality(m, vId, cnt);
                }
            }
            dvec4 vertex{pos, f};
            if(vertex.w > simplex[3].w) std::swap(simplex[3], vertex);
            if(vertex.w > simplex[2].w) std::swap(simplex[2], vertex);
            if(vertex.w > simplex[1].w) std::swap(simplex[1], vertex);
            if(vertex.w > simplex[0].w) std::swap(simplex[0], vertex);
            if((simplex[3].w - simplex[1].w) < NMValueConvergence) { terminated = true; break; }
        }
        if(terminated || (cycle >= NMSecurityCycleCount && reset))
            break;
        simplex[0] = dvec4{vo.p - dvec3(nodeShift, 0, 0), vo.w - 0};
        simplex[1] = dvec4{vo.p - dvec3(0, nodeShift, 0), vo.w - 0};
        simplex[2] = dvec4{vo.p - dvec3(0, 0, nodeShift), vo.w - 0};
        simplex[3] = vo;
        reset = true;
        cycle = 0;
    }
    pos = simplex[3].p;
    if(constrained) pos = snap(topo, pos);
}

// GradientDescentSmoother::smoothVertices (GradientDescentSmoother.cpp:73-202), constants :14-15.
static void gradientDescentVertex(Mesh& m, uint32_t vId, int64_t* cnt)
{
    const int GDSecurityCycleCount = 5;
    const double GDLocalSizeToNodeShift = 1.0 / 75.0;
    double localSize = computeLocalElementSize(m, vId);
    double nodeShift = localSize * GDLocalSizeToNodeShift;
    double originalNodeShift = nodeShift;
    const Topo& topo = m.topos[vId];
    bool constrained = isConstrained(topo);
    for(int c = 0; c < GDSecurityCycleCount; ++c)
    {
        dvec3& pos = m.verts[vId];
        double sq[6];
        dvec3 gs[6] = {
            pos + dvec3(-nodeShift, 0.0, 0.0), pos + dvec3(nodeShift, 0.0, 0.0),
            pos + dvec3(0.0, -nodeShift, 0.0), pos + dvec3(0.0, nodeShift, 0.0),
            pos + dvec3(0.0, 0.0, -nodeShift), pos + dvec3(0.0, 0.0, nodeShift)};
        if(constrained) for(int p = 0; p < 6; ++p) gs[p] = snap(topo, gs[p]);
        dvec3 originalPos = pos;
        for(int p = 0; p < 6; ++p) { pos = gs[p]; sq[p] = patchQuality(m, vId, cnt); }
        pos = originalPos;
        dvec3 gradQ(sq[1] - sq[0], sq[3] - sq[2], sq[5] - sq[4]);
        double gradQNorm = length(gradQ);
        if(gradQNorm == 0) break;
        const double OFFSETS[8] = {-0.25, 0.00, 0.10, 0.20, 0.40, 0.80, 1.20, 1.60};
        dvec3 shift = gradQ * (nodeShift / gradQNorm);
        dvec3 props[8];
        for(int p = 0; p < 8; ++p) props[p] = pos + shift * OFFSETS[p];
        if(constrained) for(int p = 0; p < 8; ++p) props[p] = snap(topo, props[p]);
        int best = 0;
        double bestQ = -std::numeric_limits<double>::infinity();
        for(int p = 0; p < 8; ++p)
        {
            pos = props[p];
            double q = patchQuality(m, vId, cnt);
            if(q > bestQ) { bestQ = q; best = p; }
        }
        pos = props[best];
        nodeShift *= std::abs(OFFSETS[best]);
        if(nodeShift < originalNodeShift / 10.0) break;
    }
}

// SpawnSearchSmoother (SpawnSearchSmoother.cpp:33-55 offsets, :161-226 search).
static void spawnSearchVertex(Mesh& m, uint32_t vId, double moveCoeff, int64_t* cnt)
{
    static dvec3 offsets[64];
    static bool init = false;
    if(!init)
    {
        int n = 0;
        for(int k = 0; k < 4; ++k) for(int j = 0; j < 4; ++j) for(int i = 0; i < 4; ++i)
            offsets[n++] = dvec3(i, j, k) - dvec3(1.5);
        offsets[0] = dvec3(0, 0, 0);
        init = true;
    }
    dvec3& pos = m.verts[vId];
    double localSize = computeLocalElementSize(m, vId);
    double scale = moveCoeff * localSize;
    const Topo& topo = m.topos[vId];
    for(int iter = 0; iter < 2; ++iter)
    {
        dvec3 props[64];
        for(int p = 0; p < 64; ++p) props[p] = pos + offsets[p] * scale;
        if(isConstrained(topo)) for(int p = 0; p < 64; ++p) props[p] = snap(topo, props[p]);
        int best = 0;
        double bestQ = -std::numeric_limits<double>::infinity();
        for(int p = 0; p < 64; ++p)
        {
            pos = props[p];
            double q = patchQuality(m, vId, cnt);
            if(q > bestQ) { bestQ = q; best = p; }
        }
        pos = props[best];
        scale /= 3.0;
    }
}

static void smoothVertex(Mesh& m, uint32_t vId, const SmoothParams& sp)
{
    int64_t* cnt = sp.evalCount ? &(*sp.evalCount)[vId] : nullptr;
    switch(sp.smoother)
    {
    case 0: nelderMeadVertex(m, vId, cnt); break;
    case 1: gradientDescentVertex(m, vId, cnt); break;
    case 2: spawnSearchVertex(m, vId, sp.spawnMoveCoeff, cnt); break;
    }
}

// Runs `fn(i)` for i in [0, n) on `threads` workers, contiguous chunks.
template<typename F>
static void parallelFor(size_t n, int threads, F fn)
{
    if(threads <= 1 || n < 64) { for(size_t i = 0; i < n; ++i) fn(i); return; }
    std::vector<std::thread> th;
    for(int t = 0; t < threads; ++t)
        th.emplace_back([&, t]() {
            size_t b = n * t / threads, e = n * (t + 1) / threads;
            for(size_t i = b; i < e; ++i) fn(i);
        });
    for(auto& t : th) t.join();
}

// ------------------------------------------------------------------------
// GETMe (GetmeSmoother.cpp:55-127, 245-374; AbstractElementWiseSmoother.cpp:53-83, 406-444).
// Element transforms are computed against the pass-start positions (the
// reference's element loop finishes before any vertex moves); each vertex
// then sums its contributions in neighborElems order, which is the order the
// reference's element loop pushes them (ascending id, tets before hexes).
// ------------------------------------------------------------------------
struct ElemOut { dvec3 vpp[8]; double weight; };

// The dual transform and volume rescale of GetmeSmoother::smoothTets (:55-100),
// before snapping.
static void getmeTetTransform(const Crew& c, const dvec3 vp[4], dvec3 vpp[4])
{
    const double lambda = 0.78;
    dvec3 n[4] = {cross(vp[3] - vp[1], vp[2] - vp[1]), cross(vp[3] - vp[2], vp[0] - vp[2]),
                  cross(vp[1] - vp[3], vp[0] - vp[3]), cross(vp[1] - vp[0], vp[2] - vp[0])};
    for(int k = 0; k < 4; ++k) vpp[k] = vp[k] + n[k] * (lambda / std::sqrt(length(n[k])));
    double volume = tetVolume(c, vp);
    double volumePrime = tetVolume(c, vpp);
    double absVolumeRation = std::abs(volume / volumePrime);
    double volumeVar = pow13(absVolumeRation);
    dvec3 center = (1.0 / 4.0) * (vp[0] + vp[1] + vp[2] + vp[3]);
    for(int k = 0; k < 4; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
}

static void getmeTet(const Mesh& m, uint32_t e, ElemOut& o)
{
    const Crew& c = m.crew;
    uint32_t vi[4]; dvec3 vp[4], vpp[4];
    for(int k = 0; k < 4; ++k) { vi[k] = m.tets[e][k]; vp[k] = m.verts[vi[k]]; }
    getmeTetTransform(c, vp, vpp);
    for(int k = 0; k < 4; ++k) if(isConstrained(m.topos[vi[k]])) vpp[k] = snap(m.topos[vi[k]], vpp[k]);
    double quality = tetQuality(c, vp);
    double qualityPrime = tetQuality(c, vpp);
    o.weight = qualityPrime / (1.0 + quality);
    for(int k = 0; k < 4; ++k) o.vpp[k] = vpp[k];
}

// GetmeSmoother::smoothHexs (:245-352), before snapping.
static void getmeHexTransform(const Crew& c, const dvec3 vp[8], dvec3 vpp[8])
{
    const double lambda = 0.78;
    dvec3 aux[6] = {
        (vp[0] + vp[1] + vp[2] + vp[3]) / 4.0, (vp[0] + vp[4] + vp[5] + vp[1]) / 4.0,
        (vp[1] + vp[5] + vp[6] + vp[2]) / 4.0, (vp[2] + vp[6] + vp[7] + vp[3]) / 4.0,
        (vp[0] + vp[3] + vp[7] + vp[4]) / 4.0, (vp[4] + vp[7] + vp[6] + vp[5]) / 4.0};
    dvec3 n[8] = {
        cross(aux[1] - aux[0], aux[4] - aux[0]), cross(aux[2] - aux[0], aux[1] - aux[0]),
        cross(aux[3] - aux[0], aux[2] - aux[0]), cross(aux[4] - aux[0], aux[3] - aux[0]),
        cross(aux[4] - aux[5], aux[1] - aux[5]), cross(aux[1] - aux[5], aux[2] - aux[5]),
        cross(aux[2] - aux[5], aux[3] - aux[5]), cross(aux[3] - aux[5], aux[4] - aux[5])};
    dvec3 bases[8] = {
        (aux[0] + aux[1] + aux[4]) / 3.0, (aux[0] + aux[2] + aux[1]) / 3.0,
        (aux[0] + aux[3] + aux[2]) / 3.0, (aux[0] + aux[4] + aux[3]) / 3.0,
        (aux[5] + aux[4] + aux[1]) / 3.0, (aux[5] + aux[1] + aux[2]) / 3.0,
        (aux[5] + aux[2] + aux[3]) / 3.0, (aux[5] + aux[3] + aux[4]) / 3.0};
    for(int k = 0; k < 8; ++k) vpp[k] = bases[k] + n[k] * (lambda / std::sqrt(length(n[k])));
    double volume = hexVolume(c, vp);
    double volumePrime = hexVolume(c, vpp);
    double absVolumeRation = std::abs(volume / volumePrime);
    double volumeVar = pow13(absVolumeRation);
    dvec3 center = (1.0 / 8.0) * (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5] + vp[6] + vp[7]);
    for(int k = 0; k < 8; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
}

static void getmeHex(const Mesh& m, uint32_t e, ElemOut& o)
{
    const Crew& c = m.crew;
    uint32_t vi[8]; dvec3 vp[8], vpp[8];
    for(int k = 0; k < 8; ++k) { vi[k] = m.hexs[e][k]; vp[k] = m.verts[vi[k]]; }
    getmeHexTransform(c, vp, vpp);
    for(int k = 0; k < 8; ++k) if(isConstrained(m.topos[vi[k]])) vpp[k] = snap(m.topos[vi[k]], vpp[k]);
    double quality = hexQuality(c, vp);
    double qualityPrime = hexQuality(c, vpp);
    o.weight = qualityPrime / (1.0 + quality);
    for(int k = 0; k < 8; ++k) o.vpp[k] = vpp[k];
}

// GetmeSmoother::smoothPris (:131-243), before snapping.
static void getmePriTransform(const Crew& c, const dvec3 vp[6], dvec3 vpp[6])
{
    const double lambda = 0.78;
    dvec3 aux[5] = {
        (vp[0] + vp[1] + vp[2]) / 3.0, (vp[0] + vp[1] + vp[4] + vp[3]) / 4.0,
        (vp[1] + vp[2] + vp[5] + vp[4]) / 4.0, (vp[2] + vp[0] + vp[3] + vp[5]) / 4.0,
        (vp[3] + vp[4] + vp[5]) / 3.0};
    dvec3 n[6] = {
        cross(aux[1] - aux[0], aux[3] - aux[0]), cross(aux[2] - aux[0], aux[1] - aux[0]),
        cross(aux[3] - aux[0], aux[2] - aux[0]), cross(aux[3] - aux[4], aux[1] - aux[4]),
        cross(aux[1] - aux[4], aux[2] - aux[4]), cross(aux[2] - aux[4], aux[3] - aux[4])};
    const double t = (4.0 / 5.0) * (1.0 - dpow(4.0 / 39.0, 0.25) * lambda);
    const double it = 1.0 - t;
    dvec3 bases[6] = {
        it * aux[0] + t * (aux[3] + aux[1]) / 2.0, it * aux[0] + t * (aux[1] + aux[2]) / 2.0,
        it * aux[0] + t * (aux[2] + aux[3]) / 2.0, it * aux[4] + t * (aux[3] + aux[1]) / 2.0,
        it * aux[4] + t * (aux[1] + aux[2]) / 2.0, it * aux[4] + t * (aux[2] + aux[3]) / 2.0};
    for(int k = 0; k < 6; ++k) vpp[k] = bases[k] + n[k] * (lambda / std::sqrt(length(n[k])));
    double volume = priVolume(c, vp);
    double volumePrime = priVolume(c, vpp);
    double absVolumeRation = std::abs(volume / volumePrime);
    double volumeVar = pow13(absVolumeRation);
    dvec3 center = (1.0 / 6.0) * (vp[0] + vp[1] + vp[2] + vp[3] + vp[4] + vp[5]);
    for(int k = 0; k < 6; ++k) vpp[k] = center + volumeVar * (vpp[k] - center);
}

static void getmePri(const Mesh& m, uint32_t e, ElemOut& o)
{
    const Crew& c = m.crew;
    uint32_t vi[6]; dvec3 vp[6], vpp[6];
    for(int k = 0; k < 6; ++k) { vi[k] = m.pris[e][k]; vp[k] = m.verts[vi[k]]; }
    getmePriTransform(c, vp, vpp);
    for(int k = 0; k < 6; ++k) if(isConstrained(m.topos[vi[k]])) vpp[k] = snap(m.topos[vi[k]], vpp[k]);
    double quality = priQuality(c, vp);
    double qualityPrime = priQuality(c, vpp);
    o.weight = qualityPrime / (1.0 + quality);
    for(int k = 0; k < 6; ++k) o.vpp[k] = vpp[k];
}

// Element-step outputs per element kind (indexed by TET, PYR, PRI, HEX).
struct GetmeOut { std::vector<ElemOut> kind[4]; };

// AbstractElementWiseSmoother::updateVertexPositions for one vertex (:406-444),
// with NotThreadSafeVertexAccum (VertexAccum.cpp:28-49) summed by pull.
static void getmeUpdateVertex(Mesh& m, uint32_t vId, const GetmeOut& out, int64_t* cnt)
{
    dvec3 posAccum(0.0);
    double weightAccum = 0.0;
    for(const NeigElem& ne : m.topos[vId].neighborElems)
    {
        const ElemOut& o = out.kind[ne.type][ne.id];
        posAccum = posAccum + o.vpp[ne.vId] * o.weight;
        weightAccum += o.weight;
    }
    dvec3 pos = m.verts[vId];
    dvec3 posPrim = pos;
    if(weightAccum != 0.0)
    {
        posPrim = posAccum / weightAccum;
        const Topo& topo = m.topos[vId];
        if(isConstrained(topo)) posPrim = snap(topo, posPrim);
        double pq = patchQuality(m, vId, cnt);
        m.verts[vId] = posPrim;
        double pqp = patchQuality(m, vId, cnt);
        if(pqp < pq) m.verts[vId] = pos;
    }
}

// ------------------------------------------------------------------------
// BATR topologist (Topologists/BatrTopologist.cpp) on meshes of tetrahedra
// whose vertices are free or fixed.
//
// The boundary algebra of those two kinds (AbstractBoundary.cpp:39-66,
// VertexConstraint.cpp:65-150, VolumeConstraint.cpp:52-60): a free vertex
// holds a VolumeConstraint, a fixed one its own VertexConstraint
// (Mesh.cpp:139-147), so
//   merge(a, b): a == b -> a; a free -> b (VolumeConstraint::merge returns its
//     argument); b free -> a (a point merged with a volume is the point); two
//     different points -> invalid (PREVENT_MERGE);
//   split(a, b): a == b -> a; otherwise a volume (point-point gives
//     SPLIT_VOLUME, point-volume the volume, volume-anything itself).
// Volume constraints are interchangeable (identity snap, same algebra).
// cureBoundaries (:1486-1712) acts on vertices of a face constraint
// (dimension 2) and on tets with three constrained vertices whose split is
// constrained; with free and fixed vertices neither exists (three distinct
// points split to the volume), so it changes nothing and is not restated.
// Meshes with sliding vertices are refused.
// ------------------------------------------------------------------------
struct BCons { int constraint = 0; dvec3 fixedPos; int fixedId = -1; bool valid = true; };

static BCons consOf(const Topo& t)
{
    BCons c;
    c.constraint = t.constraint;
    c.fixedPos = t.fixedPos;
    c.fixedId = t.fixedId;
    return c;
}
static bool sameCons(const BCons& a, const BCons& b)
{
    return a.constraint == b.constraint && (a.constraint == 0 || a.fixedId == b.fixedId);
}
static BCons mergeCons(const BCons& a, const BCons& b)
{
    if(sameCons(a, b)) return a;
    if(a.constraint == 0) return b;
    if(b.constraint == 0) return a;
    BCons bad;
    bad.valid = false;
    return bad;
}
static BCons splitCons(const BCons& a, const BCons& b)
{
    if(sameCons(a, b)) return a;
    return BCons{};
}
static void setCons(Topo& t, const BCons& c)
{
    t.constraint = c.constraint;
    t.fixedPos = c.fixedPos;
    t.fixedId = c.fixedId;
}
// (*constraint)(pos): VertexConstraint.cpp:60-63, the volume's identity.
static dvec3 applyCons(const BCons& c, const dvec3& p) { return c.constraint < 0 ? c.fixedPos : p; }

typedef std::array<uint32_t, 4> Tet4;

// AbstractEvaluator::tetQuality(mesh, sampler, measurer, tet).
static double batrQuality(const Mesh& m, const Tet4& t)
{
    const dvec3 vp[4] = {m.verts[t[0]], m.verts[t[1]], m.verts[t[2]], m.verts[t[3]]};
    return tetQuality(m.crew, vp);
}

// MetricFreeMeasurer::riemannianDistance (MetricFreeMeasurer.cpp:27-34),
// MetricWiseMeasurer::riemannianDistance (MetricWiseMeasurer.cpp:65-116).
static double batrDistance(const Crew& c, const dvec3& a, const dvec3& b)
{
    if(c.measurer == MEAS_METRIC_WISE) return riemannianDistanceWise(c, a, b);
    return length(a - b) * c.scaling;
}

// AbstractMeasurer::tetEuclideanVolume (AbstractMeasurer.cpp:61-84).
static double tetEuclideanVolume(const Mesh& m, const Tet4& t)
{
    const dvec3 *p0 = &m.verts[t[0]], *p1 = &m.verts[t[1]], *p2 = &m.verts[t[2]], *p3 = &m.verts[t[3]];
    return determinant(dmat3(*p3 - *p0, *p3 - *p1, *p3 - *p2)) / 6.0;
}

static NeigElem toTet(uint32_t id) { return NeigElem{id, TET, -1}; }
static uint32_t idOf(uint32_t x) { return x; }
static uint32_t idOf(const NeigElem& e) { return e.id; }

// BatrTopologist::popOut (:1279-1324) and make_union (:1326-1340).
template<typename T> static bool popOut(std::vector<T>& vec, uint32_t val)
{
    const size_t n = vec.size();
    for(size_t i = 0; i < n; ++i)
        if(idOf(vec[i]) == val)
        {
            std::swap(vec[i], vec.back());
            vec.pop_back();
            return true;
        }
    return false;
}
template<typename T> static bool popOutAll(std::vector<T>& vec, const std::vector<uint32_t>& vals)
{
    size_t popCount = 0, i = 0;
    while(i < vec.size())
    {
        bool found = false;
        for(uint32_t val : vals)
            if(idOf(vec[i]) == val)
            {
                std::swap(vec[i], vec.back());
                vec.pop_back();
                found = true;
                ++popCount;
                break;
            }
        if(!found) ++i;
    }
    return popCount == vals.size();
}
static bool makeUnion(std::vector<uint32_t>& vec, uint32_t val)
{
    for(uint32_t x : vec)
        if(x == val) return false;
    vec.push_back(val);
    return true;
}

// BatrTopologist::buildVertNeighborhood (:1342-1357).
static void buildVertNeighborhood(Mesh& m, uint32_t vId)
{
    std::set<uint32_t> s;
    for(const NeigElem& e : m.topos[vId].neighborElems)
    {
        const Tet4& t = m.tets[e.id];
        for(int k = 0; k < 4; ++k)
            if(t[k] != vId) s.insert(t[k]);
    }
    m.topos[vId].neighborVerts.assign(s.begin(), s.end());
}

// BatrTopologist::findRing (:1359-1393).
static void findRing(const Mesh& m, uint32_t vId, uint32_t nId, std::vector<uint32_t>& ringVerts,
                     std::vector<uint32_t>& ringElems)
{
    for(const NeigElem& ve : m.topos[vId].neighborElems)
        for(const NeigElem& ne : m.topos[nId].neighborElems)
            if(ve.id == ne.id)
            {
                ringElems.push_back(ve.id);
                break;
            }
    std::set<uint32_t> s;
    for(uint32_t e : ringElems)
        for(int k = 0; k < 4; ++k)
            if(m.tets[e][k] != vId && m.tets[e][k] != nId) s.insert(m.tets[e][k]);
    ringVerts.assign(s.begin(), s.end());
}

// BatrTopologist::findExclusiveElems (:1395-1418).
static void findExclusiveElems(const Mesh& m, uint32_t vId, const std::vector<uint32_t>& ringElems,
                               std::vector<uint32_t>& exElems)
{
    for(const NeigElem& ve : m.topos[vId].neighborElems)
    {
        bool in = false;
        for(uint32_t r : ringElems)
            if(ve.id == r)
            {
                in = true;
                break;
            }
        if(!in) exElems.push_back(ve.id);
    }
}

// BatrTopologist::trimVerts (:1420-1455).
static void trimVerts(Mesh& m, const std::vector<bool>& aliveVerts)
{
    size_t copy = 0;
    for(size_t vId = 0; vId < m.verts.size(); ++vId)
    {
        if(!aliveVerts[vId]) continue;
        if(copy != vId)
        {
            for(uint32_t nv : m.topos[vId].neighborVerts)
                for(uint32_t& nnv : m.topos[nv].neighborVerts)
                    if(nnv == vId)
                    {
                        nnv = (uint32_t)copy;
                        break;
                    }
            for(const NeigElem& ne : m.topos[vId].neighborElems)
            {
                Tet4& t = m.tets[ne.id];
                if(t[0] == vId) t[0] = (uint32_t)copy;
                else if(t[1] == vId) t[1] = (uint32_t)copy;
                else if(t[2] == vId) t[2] = (uint32_t)copy;
                else if(t[3] == vId) t[3] = (uint32_t)copy;
            }
            m.verts[copy] = m.verts[vId];
            m.topos[copy] = m.topos[vId];
        }
        ++copy;
    }
    m.verts.resize(copy);
    m.topos.resize(copy);
}

// BatrTopologist::trimTets (:1457-1484).
static void trimTets(Mesh& m, const std::vector<bool>& aliveTets)
{
    size_t copy = 0;
    const size_t n = m.tets.size();
    for(size_t tId = 0; tId < n; ++tId)
    {
        if(!aliveTets[tId]) continue;
        if(copy != tId)
        {
            const Tet4 t = m.tets[tId];
            for(int i = 0; i < 4; ++i)
                for(NeigElem& ne : m.topos[t[i]].neighborElems)
                    if(ne.id == tId)
                    {
                        ne.id = (uint32_t)copy;
                        break;
                    }
            m.tets[copy] = t;
        }
        ++copy;
    }
    m.tets.resize(copy);
}

// MeshTet::tris (Mesh.cpp:30-35): the face opposite each corner.
static const int TET_TRIS[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};

// BatrTopologist::faceSwapping (:635-836).
static size_t faceSwapping(Mesh& m)
{
    std::vector<Tet4>& tets = m.tets;
    std::vector<Topo>& topos = m.topos;
    size_t faceSwapCount = 0;
    bool stillTetsToTry = true;
    std::vector<bool> tetsToTry(tets.size(), true);
    while(stillTetsToTry)
    {
        stillTetsToTry = false;
        for(size_t t = 0; t < tets.size(); ++t)
        {
            if(!tetsToTry[t]) continue;
            tetsToTry[t] = false;
            const Tet4 tet = tets[t];
            for(int f = 0; f < 4; ++f)
            {
                const uint32_t tri[3] = {tet[TET_TRIS[f][0]], tet[TET_TRIS[f][1]], tet[TET_TRIS[f][2]]};
                if(!((tri[0] < tri[1] && tri[1] < tri[2]) || (tri[1] < tri[2] && tri[2] < tri[0]) ||
                     (tri[2] < tri[0] && tri[0] < tri[1])))
                    continue;
                uint32_t nb = 0, common[2] = {0, 0};
                std::vector<NeigElem>& n0 = topos[tri[0]].neighborElems;
                std::vector<NeigElem>& n1 = topos[tri[1]].neighborElems;
                std::vector<NeigElem>& n2 = topos[tri[2]].neighborElems;
                for(const NeigElem& a : n0)
                {
                    bool in = false;
                    for(const NeigElem& b : n1)
                        if(a.id == b.id)
                        {
                            in = true;
                            break;
                        }
                    if(!in) continue;
                    in = false;
                    for(const NeigElem& c : n2)
                        if(a.id == c.id)
                        {
                            in = true;
                            break;
                        }
                    if(in)
                    {
                        if(nb < 2) common[nb] = a.id;
                        ++nb;
                    }
                }
                if(nb != 2) continue;

                const uint32_t tOp = tet[f];
                const uint32_t nt = ((uint32_t)t != common[0]) ? common[0] : common[1];
                const Tet4 ntet = tets[nt];
                uint32_t nOp = ntet[0];
                if(ntet[1] != tri[0] && ntet[1] != tri[1] && ntet[1] != tri[2]) nOp = ntet[1];
                else if(ntet[2] != tri[0] && ntet[2] != tri[1] && ntet[2] != tri[2]) nOp = ntet[2];
                else if(ntet[3] != tri[0] && ntet[3] != tri[1] && ntet[3] != tri[2]) nOp = ntet[3];

                const dvec3 vPos = m.verts[tOp];
                const dvec3 nPos = m.verts[nOp];
                const dvec3 v0 = m.verts[tri[0]] - vPos;
                const dvec3 v1 = m.verts[tri[1]] - vPos;
                const dvec3 v2 = m.verts[tri[2]] - vPos;
                const dvec3 vn = nPos - vPos;
                if(dot(cross(v0, v1), vn) < 0.0 || dot(cross(v1, v2), vn) < 0.0 || dot(cross(v2, v0), vn) < 0.0)
                    continue;

                const double tQual = batrQuality(m, tet);
                const double nQual = batrQuality(m, ntet);
                const double hQual = 2.0 / (1.0 / tQual + 1.0 / nQual);
                const Tet4 nt0 = {tOp, tri[0], tri[1], nOp};
                const Tet4 nt1 = {tOp, tri[1], tri[2], nOp};
                const Tet4 nt2 = {tOp, tri[2], tri[0], nOp};
                const double q0 = batrQuality(m, nt0);
                const double q1 = batrQuality(m, nt1);
                const double q2 = batrQuality(m, nt2);
                const double newHQual = 3.0 / (1 / q0 + 1 / q1 + 1 / q2);
                if(newHQual <= hQual) continue;

                tets[t] = nt0;
                tets[nt] = nt1;
                const uint32_t lt = (uint32_t)tets.size();
                tets.push_back(nt2);
                topos[tOp].neighborElems.push_back(toTet(nt));
                topos[tOp].neighborElems.push_back(toTet(lt));
                topos[tOp].neighborVerts.push_back(nOp);
                topos[nOp].neighborElems.push_back(toTet((uint32_t)t));
                topos[nOp].neighborElems.push_back(toTet(lt));
                topos[nOp].neighborVerts.push_back(tOp);
                for(NeigElem& e : n0)
                    if(e.id == nt)
                    {
                        e.id = lt;
                        break;
                    }
                for(NeigElem& e : n2)
                    if(e.id == t)
                    {
                        e.id = lt;
                        break;
                    }
                stillTetsToTry = true;
                tetsToTry.push_back(true);
                tetsToTry[nt] = true;
                tetsToTry[t] = true;
                ++faceSwapCount;
                break;
            }
        }
    }
    return faceSwapCount;
}

// BatrTopologist's ring configurations (BatrTopologist.cpp:36-60): per ring
// size, the triangulations of the ring polygon and how many rotations of
// each are distinct.
struct RingConfig { uint32_t rotCount; std::vector<std::array<uint32_t, 3>> tris; };
static const std::vector<std::vector<RingConfig>>& ringConfigDictionary()
{
    static const std::vector<std::vector<RingConfig>> d = {
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
    return d;
}

// BatrTopologist::edgeSwapping (:838-1278).
static size_t edgeSwapping(Mesh& m)
{
    std::vector<Tet4>& tets = m.tets;
    std::vector<Topo>& topos = m.topos;
    const std::vector<std::vector<RingConfig>>& dict = ringConfigDictionary();
    bool stillVertsToTry = true;
    std::vector<bool> vertsToTry(m.verts.size(), true);
    std::vector<bool> aliveVerts(m.verts.size(), true);
    std::vector<bool> aliveTets(tets.size(), true);
    std::vector<uint32_t> deadTets;
    size_t total = 0;

    while(stillVertsToTry)
    {
        stillVertsToTry = false;
        for(size_t vId = 0; vId < m.verts.size(); ++vId)
        {
            if(!vertsToTry[vId]) continue;
            vertsToTry[vId] = false;
            Topo& vTopo = topos[vId];
            std::vector<uint32_t>& vVerts = vTopo.neighborVerts;
            std::vector<NeigElem>& vElems = vTopo.neighborElems;

            for(size_t nAr = 0; nAr < vVerts.size(); ++nAr)
            {
                const uint32_t nId = vVerts[nAr];
                if(nId < vId) continue;

                std::vector<uint32_t> ringVerts, ringElems;
                findRing(m, (uint32_t)vId, nId, ringVerts, ringElems);

                // Order the ring by walking its tets (:889-962).
                std::vector<uint32_t> elemPool = ringElems;
                int notFoundCount = 0;
                size_t r = 0, nx = 1;
                while(nx != ringVerts.size())
                {
                    bool found = false;
                    for(size_t i = 0; i < elemPool.size() && !found; ++i)
                    {
                        const Tet4 tet = tets[elemPool[i]];
                        for(int j = 0; j < 4 && !found; ++j)
                        {
                            if(tet[j] != ringVerts[r]) continue;
                            for(int k = 0; k < 4 && !found; ++k)
                            {
                                if(k == j || tet[k] == vId || tet[k] == nId) continue;
                                for(size_t w = nx; w != ringVerts.size() && !found; ++w)
                                    if(ringVerts[w] == tet[k])
                                    {
                                        std::swap(elemPool[i], elemPool.back());
                                        elemPool.pop_back();
                                        std::swap(ringVerts[w], ringVerts[nx]);
                                        found = true;
                                    }
                            }
                        }
                    }
                    if(!found)
                    {
                        if(++notFoundCount >= 2) break;
                        size_t cb = 0, cf = r;
                        while(cb < cf) std::swap(ringVerts[cb++], ringVerts[cf--]);
                    }
                    else
                    {
                        ++r;
                        ++nx;
                    }
                }
                if(nx != ringVerts.size())
                {
                    // A ring that cannot be closed ends the operation (:964-993).
                    trimTets(m, aliveTets);
                    return 0;
                }

                const size_t ringSize = ringVerts.size();
                if(ringSize < 3 || ringSize >= dict.size()) continue;

                // Counter-clockwise around v-n (:1009-1029); GLM's dvec3
                // default constructor zero-fills.
                dvec3 ringAir(0.0, 0.0, 0.0);
                const dvec3 ring0 = m.verts[ringVerts[0]];
                for(size_t i = 2; i < ringSize; ++i)
                    ringAir = ringAir + cross(m.verts[ringVerts[i - 1]] - ring0, m.verts[ringVerts[i]] - ring0);
                const dvec3 vPos = m.verts[vId];
                const dvec3 nPos = m.verts[nId];
                const dvec3 vn = nPos - vPos;
                if(dot(ringAir, vn) < 0.0)
                {
                    size_t i = 0, j = ringVerts.size() - 1;
                    while(i < j) std::swap(ringVerts[i++], ringVerts[j--]);
                }

                double minQual = INFINITY;
                for(uint32_t e : ringElems) minQual = gmin(minQual, batrQuality(m, tets[e]));

                // Outsider check (:1041-1061): every ring edge must turn the
                // same way around v-n.
                bool outsidersOk = true;
                for(size_t i = 1; i <= ringSize; ++i)
                {
                    const uint32_t aId = ringVerts[i - 1];
                    const uint32_t bId = ringVerts[i % ringSize];
                    const dvec3 normal = cross(m.verts[aId] - vPos, m.verts[bId] - vPos);
                    if(dot(vn, normal) < 0.0)
                    {
                        outsidersOk = false;
                        break;
                    }
                }
                if(!outsidersOk) continue;

                int bestConf = -1, bestRot = -1;
                const std::vector<RingConfig>& confs = dict[ringSize];
                for(size_t conf = 0; conf < confs.size(); ++conf)
                    for(uint32_t rot = 0; rot < confs[conf].rotCount; ++rot)
                    {
                        double rotMin = INFINITY;
                        for(const auto& rt : confs[conf].tris)
                        {
                            const uint32_t a = ringVerts[(rt[0] + rot) % ringSize];
                            const uint32_t b = ringVerts[(rt[1] + rot) % ringSize];
                            const uint32_t c = ringVerts[(rt[2] + rot) % ringSize];
                            const double qv = batrQuality(m, Tet4{b, a, c, (uint32_t)vId});
                            rotMin = gmin(rotMin, qv);
                            if(qv < minQual) break;
                            const double qn = batrQuality(m, Tet4{a, b, c, nId});
                            rotMin = gmin(rotMin, qn);
                            if(qn < minQual) break;
                        }
                        if(rotMin > minQual)
                        {
                            bestConf = (int)conf;
                            bestRot = (int)rot;
                            minQual = rotMin;
                        }
                    }
                if(bestConf < 0) continue;
                ++total;

                // Topology update (:1122-1212).
                Topo& nTopo = topos[nId];
                popOutAll(vElems, ringElems);
                popOutAll(nTopo.neighborElems, ringElems);
                popOut(vVerts, nId);
                popOut(nTopo.neighborVerts, (uint32_t)vId);
                for(uint32_t rv : ringVerts) popOutAll(topos[rv].neighborElems, ringElems);
                for(uint32_t e : ringElems)
                {
                    aliveTets[e] = false;
                    deadTets.push_back(e);
                }
                std::vector<Tet4> bestTets;
                for(const auto& rt : confs[bestConf].tris)
                {
                    const uint32_t a = ringVerts[(rt[0] + bestRot) % ringSize];
                    const uint32_t b = ringVerts[(rt[1] + bestRot) % ringSize];
                    const uint32_t c = ringVerts[(rt[2] + bestRot) % ringSize];
                    bestTets.push_back(Tet4{b, a, c, (uint32_t)vId});
                    bestTets.push_back(Tet4{a, b, c, nId});
                    makeUnion(topos[a].neighborVerts, b);
                    makeUnion(topos[a].neighborVerts, c);
                    makeUnion(topos[b].neighborVerts, a);
                    makeUnion(topos[b].neighborVerts, c);
                    makeUnion(topos[c].neighborVerts, a);
                    makeUnion(topos[c].neighborVerts, b);
                }
                for(const Tet4& t : bestTets)
                {
                    uint32_t id;
                    if(deadTets.empty())
                    {
                        id = (uint32_t)tets.size();
                        tets.push_back(t);
                        aliveTets.push_back(true);
                    }
                    else
                    {
                        id = deadTets.back();
                        deadTets.pop_back();
                        tets[id] = t;
                        aliveTets[id] = true;
                    }
                    for(int k = 0; k < 4; ++k) topos[t[k]].neighborElems.push_back(toTet(id));
                }
                stillVertsToTry = true;
                for(uint32_t x : vTopo.neighborVerts) vertsToTry[x] = true;
                for(uint32_t x : nTopo.neighborVerts) vertsToTry[x] = true;
            }
        }
    }
    trimTets(m, aliveTets);
    trimVerts(m, aliveVerts);
    return total;
}

// BatrTopologist::edgeSplitMerge (:137-633); edge lengths in the crew's
// metric (AbstractTopologist's minEdgeLength / maxEdgeLength).
static void edgeSplitMerge(Mesh& m, int refinementSweeps, double minEdge, double maxEdge, int64_t* splits,
                           int64_t* merges)
{
    std::vector<Tet4>& tets = m.tets;
    std::vector<Topo>& topos = m.topos;
    const double MIN_GEN_QUALITY = 1e-6;   // _minAcceptableGenQuality (:38)
    bool stillVertsToTry = true;
    std::vector<bool> aliveVerts(m.verts.size(), true);
    std::vector<bool> vertsToTry(m.verts.size(), true);
    std::vector<bool> aliveTets(tets.size(), true);
    std::vector<uint32_t> deadVerts, deadTets;

    for(int pass = 0; pass < refinementSweeps && stillVertsToTry; ++pass)
    {
        stillVertsToTry = false;
        for(size_t vId = 0; vId < m.verts.size(); ++vId)
        {
            if(!aliveVerts[vId]) continue;
            vertsToTry[vId] = false;

            uint32_t nId = (uint32_t)-1;
            double dist = INFINITY, priority = 0.0;
            BCons constraint;
            bool have = false;
            for(size_t nAr = 0; nAr < topos[vId].neighborVerts.size(); ++nAr)
            {
                const uint32_t cand = topos[vId].neighborVerts[nAr];
                if(vId > cand && !vertsToTry[cand]) continue;
                const double cd = batrDistance(m.crew, m.verts[vId], m.verts[cand]);
                if(cd < minEdge)
                {
                    const double cp = minEdge / cd;
                    if(cp > priority)
                    {
                        const BCons mc = mergeCons(consOf(topos[vId]), consOf(topos[cand]));
                        if(mc.valid)
                        {
                            constraint = mc;
                            have = true;
                            priority = cp;
                            dist = cd;
                            nId = cand;
                        }
                    }
                }
                else if(cd > maxEdge)
                {
                    const double cp = cd / maxEdge;
                    if(cp > priority)
                    {
                        constraint = splitCons(consOf(topos[vId]), consOf(topos[cand]));
                        have = true;
                        priority = cp;
                        dist = cd;
                        nId = cand;
                    }
                }
            }
            if(!have) continue;

            const dvec3 vert = m.verts[vId];
            const dvec3 neig = m.verts[nId];
            std::vector<uint32_t> ringVerts, ringElems, nExElems;
            findRing(m, (uint32_t)vId, nId, ringVerts, ringElems);
            findExclusiveElems(m, nId, ringElems, nExElems);

            if(dist < minEdge)
            {
                // Merge n into v (:257-400).
                std::vector<uint32_t> vExElems;
                findExclusiveElems(m, (uint32_t)vId, ringElems, vExElems);
                std::vector<uint32_t> allEx(vExElems);
                allEx.insert(allEx.end(), nExElems.begin(), nExElems.end());
                dvec3 middle = (vert + neig) / 2.0;
                middle = applyCons(constraint, middle);
                double minQuality = 1.0;
                for(uint32_t e : allEx) minQuality = gmin(minQuality, batrQuality(m, tets[e]));
                minQuality = gmax(minQuality / priority, MIN_GEN_QUALITY);
                m.verts[vId] = m.verts[nId] = middle;
                bool conformal = true;
                for(uint32_t e : allEx)
                    if(batrQuality(m, tets[e]) <= minQuality)
                    {
                        conformal = false;
                        break;
                    }
                if(!conformal)
                {
                    m.verts[vId] = vert;
                    m.verts[nId] = neig;
                    continue;
                }
                deadVerts.push_back(nId);
                aliveVerts[nId] = false;
                for(uint32_t e : ringElems)
                {
                    deadTets.push_back(e);
                    aliveTets[e] = false;
                }
                for(uint32_t e : nExElems)
                {
                    Tet4& t = tets[e];
                    if(t[0] == nId) t[0] = (uint32_t)vId;
                    else if(t[1] == nId) t[1] = (uint32_t)vId;
                    else if(t[2] == nId) t[2] = (uint32_t)vId;
                    else if(t[3] == nId) t[3] = (uint32_t)vId;
                }
                const std::vector<uint32_t> ringCopy = ringVerts;
                for(uint32_t rv : ringCopy)
                {
                    popOutAll(topos[rv].neighborElems, ringElems);
                    if(topos[rv].neighborElems.empty())
                    {
                        popOut(ringVerts, rv);
                        aliveVerts[rv] = false;
                        deadVerts.push_back(rv);
                    }
                }
                Topo& vTopo = topos[vId];
                setCons(vTopo, constraint);
                vTopo.neighborElems.clear();
                for(uint32_t e : vExElems) vTopo.neighborElems.push_back(toTet(e));
                for(uint32_t e : nExElems) vTopo.neighborElems.push_back(toTet(e));
                buildVertNeighborhood(m, (uint32_t)vId);
                for(uint32_t x : std::vector<uint32_t>(vTopo.neighborVerts)) buildVertNeighborhood(m, x);
                ++*merges;
                stillVertsToTry = true;
                vertsToTry[vId] = true;
                vertsToTry[nId] = false;
                for(uint32_t x : vTopo.neighborVerts) vertsToTry[x] = true;
            }
            else if(dist > maxEdge)
            {
                // Split v-n at w (:402-590).
                dvec3 middle = (vert + neig) / 2.0;
                middle = applyCons(constraint, middle);
                Topo wTopo;
                setCons(wTopo, constraint);
                uint32_t wId;
                if(deadVerts.empty())
                {
                    wId = (uint32_t)m.verts.size();
                    topos.push_back(wTopo);
                    m.verts.push_back(middle);
                    aliveVerts.push_back(true);
                    vertsToTry.push_back(true);
                }
                else
                {
                    wId = deadVerts.back();
                    deadVerts.pop_back();
                    topos[wId] = wTopo;
                    m.verts[wId] = middle;
                    aliveVerts[wId] = true;
                    vertsToTry[wId] = true;
                }
                double minQuality = 1.0;
                for(uint32_t e : ringElems) minQuality = gmin(minQuality, batrQuality(m, tets[e]));
                minQuality = gmax(minQuality / priority, MIN_GEN_QUALITY);
                auto halves = [&](const Tet4& t, Tet4& vT, Tet4& nT) {
                    vT = t;
                    nT = t;
                    for(int k = 0; k < 4; ++k)
                    {
                        if(t[k] == vId) nT[k] = wId;
                        else if(t[k] == nId) vT[k] = wId;
                    }
                };
                bool conformal = true;
                for(uint32_t e : ringElems)
                {
                    Tet4 vT, nT;
                    halves(tets[e], vT, nT);
                    if(batrQuality(m, vT) <= minQuality || batrQuality(m, nT) <= minQuality)
                    {
                        conformal = false;
                        break;
                    }
                }
                if(!conformal)
                {
                    vertsToTry[wId] = false;
                    aliveVerts[wId] = false;
                    deadVerts.push_back(wId);
                    continue;
                }
                for(uint32_t e : ringElems)
                {
                    const Tet4 t = tets[e];
                    for(int k = 0; k < 4; ++k) popOut(topos[t[k]].neighborElems, e);
                    deadTets.push_back(e);
                    aliveTets[e] = false;
                    Tet4 vT, nT;
                    halves(t, vT, nT);
                    if(tetEuclideanVolume(m, vT) < 0.0) std::swap(vT[0], vT[1]);
                    if(tetEuclideanVolume(m, nT) < 0.0) std::swap(nT[0], nT[1]);
                    uint32_t ids[2];
                    const Tet4* halvesOut[2] = {&vT, &nT};
                    for(int h = 0; h < 2; ++h)
                    {
                        if(deadTets.empty())
                        {
                            ids[h] = (uint32_t)tets.size();
                            aliveTets.push_back(true);
                            tets.push_back(*halvesOut[h]);
                        }
                        else
                        {
                            ids[h] = deadTets.back();
                            deadTets.pop_back();
                            aliveTets[ids[h]] = true;
                            tets[ids[h]] = *halvesOut[h];
                        }
                    }
                    for(int k = 0; k < 4; ++k) topos[vT[k]].neighborElems.push_back(toTet(ids[0]));
                    for(int k = 0; k < 4; ++k) topos[nT[k]].neighborElems.push_back(toTet(ids[1]));
                }
                buildVertNeighborhood(m, wId);
                for(uint32_t x : std::vector<uint32_t>(topos[wId].neighborVerts))
                {
                    buildVertNeighborhood(m, x);
                    vertsToTry[x] = true;
                }
                ++*splits;
                stillVertsToTry = true;
            }
        }
    }
    trimTets(m, aliveTets);
    trimVerts(m, aliveVerts);
}

// BatrTopologist::restructureMesh (:76-128) followed by its
// Mesh::compileTopology (compileNeighborhoods + NodeGroups::build).
// counts: face swaps, edge swaps, splits, merges, passes.  Returns 0, or -2
// for a mesh with sliding vertices.  Meshes with prisms or hexahedra, or
// without tetrahedra, are left as they are (needTopologicalModifications,
// :67-74).
static int restructureMesh(Mesh& m, int topoPasses, int refinementSweeps, double minEdge, double maxEdge,
                           int64_t counts[5])
{
    for(int k = 0; k < 5; ++k) counts[k] = 0;
    if(m.tets.empty() || !(m.pris.empty() && m.hexs.empty())) return 0;
    for(const Topo& t : m.topos)
        if(t.constraint > 0) return -2;
    int64_t lastOps = 0;
    int64_t passDone = 0;
    while(true)
    {
        const int64_t fs = (int64_t)faceSwapping(m);
        const int64_t es = (int64_t)edgeSwapping(m);
        int64_t sp = 0, mg = 0;
        edgeSplitMerge(m, refinementSweeps, minEdge, maxEdge, &sp, &mg);
        counts[0] += fs;
        counts[1] += es;
        counts[2] += sp;
        counts[3] += mg;
        const int64_t ops = fs + es + sp + mg;
        ++passDone;
        if(ops == 0 || ops == lastOps || passDone >= topoPasses) break;
        lastOps = ops;
    }
    counts[4] = passDone;
    compileNeighborhoods(m);
    buildNodeGroups(m);
    return 0;
}

} // namespace orc

using namespace orc;

// ------------------------------------------------------------------------
// C ABI for the tests (ctypes).
// ------------------------------------------------------------------------
extern "C" {

void* orc_mesh_create_full(int64_t nv, const double* pos, const int32_t* fixed, const int32_t* cons,
                           int64_t ncons, const int32_t* kinds, const double* params, int64_t ntet,
                           const uint32_t* tets, int64_t npri, const uint32_t* pris, int64_t nhex,
                           const uint32_t* hexs);

void* orc_mesh_create(int64_t nv, const double* pos, const int32_t* fixed,
                      int64_t ntet, const uint32_t* tets, int64_t nhex, const uint32_t* hexs)
{
    return orc_mesh_create_full(nv, pos, fixed, nullptr, 0, nullptr, nullptr, ntet, tets, 0, nullptr, nhex, hexs);
}

// A mesh with sliding constraints: cons[v] = index into the table or -1;
// kinds[k] 1 plane (params a = point, b = normal), 2 segment (a, b ends).
// One snap of `pos` onto a constraint built as orc_mesh_create_constrained
// builds it (kind 1 plane through a with normal b, 2 segment a-b).
void orc_constraint_snap(int kind, const double* a, const double* b, const double* pos, double* out)
{
    SlideConstraint c;
    c.kind = kind;
    if(kind == 1)
    {
        c.p = dvec3(a[0], a[1], a[2]);
        c.n = normalize(dvec3(b[0], b[1], b[2]));
    }
    else
    {
        c.a = dvec3(a[0], a[1], a[2]);
        c.b = dvec3(b[0], b[1], b[2]);
        c.u = normalize(c.b - c.a);
        c.length = length(c.b - c.a);
    }
    Topo t;
    t.constraint = 1;
    t.slide = &c;
    const dvec3 r = snap(t, dvec3(pos[0], pos[1], pos[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

void* orc_mesh_create_constrained(int64_t nv, const double* pos, const int32_t* fixed, const int32_t* cons,
                                  int64_t ncons, const int32_t* kinds, const double* params, int64_t ntet,
                                  const uint32_t* tets, int64_t nhex, const uint32_t* hexs)
{
    return orc_mesh_create_full(nv, pos, fixed, cons, ncons, kinds, params, ntet, tets, 0, nullptr, nhex, hexs);
}

// A mesh of tetrahedra, prisms (6 vertices each, MeshPri) and hexahedra,
// with optional sliding constraints as orc_mesh_create_constrained.
void* orc_mesh_create_full(int64_t nv, const double* pos, const int32_t* fixed, const int32_t* cons,
                           int64_t ncons, const int32_t* kinds, const double* params, int64_t ntet,
                           const uint32_t* tets, int64_t npri, const uint32_t* pris, int64_t nhex,
                           const uint32_t* hexs)
{
    Mesh* m = new Mesh();
    m->slides.resize(ncons);
    for(int64_t k = 0; k < ncons; ++k)
    {
        SlideConstraint& c = m->slides[k];
        const double* q = params + 6 * k;
        c.kind = kinds[k];
        if(c.kind == 1)
        {
            c.p = dvec3(q[0], q[1], q[2]);
            c.n = normalize(dvec3(q[3], q[4], q[5]));
        }
        else
        {
            c.a = dvec3(q[0], q[1], q[2]);
            c.b = dvec3(q[3], q[4], q[5]);
            c.u = normalize(c.b - c.a);
            c.length = length(c.b - c.a);
        }
    }
    m->verts.resize(nv);
    m->topos.resize(nv);
    for(int64_t v = 0; v < nv; ++v)
    {
        m->verts[v] = dvec3(pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]);
        m->topos[v].fixedPos = m->verts[v];
        if(fixed[v])
        {
            m->topos[v].constraint = -1;
            m->topos[v].fixedId = (int)v;
        }
        else if(cons && cons[v] >= 0)
        {
            m->topos[v].constraint = cons[v] + 1;
            m->topos[v].slide = &m->slides[cons[v]];
        }
        else m->topos[v].constraint = 0;
    }
    m->tets.resize(ntet);
    for(int64_t e = 0; e < ntet; ++e) for(int k = 0; k < 4; ++k) m->tets[e][k] = tets[4 * e + k];
    m->pris.resize(npri);
    for(int64_t e = 0; e < npri; ++e) for(int k = 0; k < 6; ++k) m->pris[e][k] = pris[6 * e + k];
    m->hexs.resize(nhex);
    for(int64_t e = 0; e < nhex; ++e) for(int k = 0; k < 8; ++k) m->hexs[e][k] = hexs[8 * e + k];
    compileNeighborhoods(*m);
    buildNodeGroups(*m);
    samplerRotation(m->crew);
    return m;
}

void orc_mesh_destroy(void* h) { delete (Mesh*)h; }

int orc_set_libm(int mode) { int old = g_libm; g_libm = mode; return old; }

int64_t orc_mesh_nverts(void* h) { return (int64_t)((Mesh*)h)->verts.size(); }

void orc_mesh_positions(void* h, double* out)
{
    Mesh* m = (Mesh*)h;
    for(size_t v = 0; v < m->verts.size(); ++v) { out[3*v] = m->verts[v].x; out[3*v+1] = m->verts[v].y; out[3*v+2] = m->verts[v].z; }
}

void orc_mesh_set_positions(void* h, const double* in)
{
    Mesh* m = (Mesh*)h;
    for(size_t v = 0; v < m->verts.size(); ++v) m->verts[v] = dvec3(in[3*v], in[3*v+1], in[3*v+2]);
}

void orc_mesh_permutation(void* h, int64_t* out)
{
    Mesh* m = (Mesh*)h;
    for(size_t v = 0; v < m->permutation.size(); ++v) out[v] = m->permutation[v];
}

void orc_mesh_elements(void* h, uint32_t* tets, uint32_t* hexs)
{
    Mesh* m = (Mesh*)h;
    for(size_t e = 0; e < m->tets.size(); ++e) for(int k = 0; k < 4; ++k) tets[4*e+k] = m->tets[e][k];
    for(size_t e = 0; e < m->hexs.size(); ++e) for(int k = 0; k < 8; ++k) hexs[8*e+k] = m->hexs[e][k];
}

int64_t orc_mesh_npris(void* h) { return (int64_t)((Mesh*)h)->pris.size(); }

int64_t orc_mesh_ntets(void* h) { return (int64_t)((Mesh*)h)->tets.size(); }

// Fixed flags per vertex (1 fixed, 0 otherwise).
void orc_mesh_fixed(void* h, int32_t* out)
{
    const Mesh& m = *(Mesh*)h;
    for(size_t v = 0; v < m.topos.size(); ++v) out[v] = m.topos[v].constraint < 0 ? 1 : 0;
}

// BatrTopologist::restructureMesh with Schedule::topoOperationPassCount =
// topoPasses and Schedule::refinementSweepCount = refinementSweeps, edge
// lengths AbstractTopologist::minEdgeLength / maxEdgeLength; the mesh is then
// recompiled (neighbourhoods and node groups, renumbered).  counts: face
// swaps, edge swaps, splits, merges, passes.  Returns 0, or -2 when a vertex
// slides on a constraint.
int orc_mesh_restructure(void* h, int topoPasses, int refinementSweeps, double minEdge, double maxEdge,
                         int64_t* counts)
{
    return restructureMesh(*(Mesh*)h, topoPasses, refinementSweeps, minEdge, maxEdge, counts);
}


void orc_mesh_prisms(void* h, uint32_t* pris)
{
    Mesh* m = (Mesh*)h;
    for(size_t e = 0; e < m->pris.size(); ++e) for(int k = 0; k < 6; ++k) pris[6*e+k] = m->pris[e][k];
}

// Sizes of the topology arrays: [sum neighborElems, sum neighborVerts].
void orc_mesh_topology_sizes(void* h, int64_t* out)
{
    Mesh* m = (Mesh*)h;
    int64_t ne = 0, nn = 0;
    for(const Topo& t : m->topos) { ne += t.neighborElems.size(); nn += t.neighborVerts.size(); }
    out[0] = ne; out[1] = nn;
}

// CSR export; elem_ref packs (id << 5) | (type << 3) | vId.
void orc_mesh_topology(void* h, int64_t* elem_off, uint64_t* elem_ref, int64_t* nbr_off, uint32_t* nbr)
{
    Mesh* m = (Mesh*)h;
    int64_t a = 0, b = 0;
    for(size_t v = 0; v < m->topos.size(); ++v)
    {
        elem_off[v] = a; nbr_off[v] = b;
        for(const NeigElem& ne : m->topos[v].neighborElems) elem_ref[a++] = ((uint64_t)ne.id << 5) | ((uint64_t)ne.type << 3) | (uint64_t)ne.vId;
        for(uint32_t n : m->topos[v].neighborVerts) nbr[b++] = n;
    }
    elem_off[m->topos.size()] = a; nbr_off[m->topos.size()] = b;
}

int64_t orc_mesh_ngroups(void* h) { return (int64_t)((Mesh*)h)->parallelGroups.size(); }

// Per group: [boundary.begin, boundary.end, subsurface.begin, subsurface.end,
// interior.begin, interior.end]; then node ranges [fixed, bound, subsu, inter].
void orc_mesh_groups(void* h, int64_t* ranges, int64_t* typeRanges)
{
    Mesh* m = (Mesh*)h;
    for(size_t g = 0; g < m->parallelGroups.size(); ++g)
    {
        const ParallelGroup& pg = m->parallelGroups[g];
        int64_t* r = ranges + 6 * g;
        r[0] = pg.boundaryRange.begin; r[1] = pg.boundaryRange.end;
        r[2] = pg.subsurfaceRange.begin; r[3] = pg.subsurfaceRange.end;
        r[4] = pg.interiorRange.begin; r[5] = pg.interiorRange.end;
    }
    typeRanges[0] = m->fixedNodes.begin; typeRanges[1] = m->fixedNodes.end;
    typeRanges[2] = m->boundaryNodes.begin; typeRanges[3] = m->boundaryNodes.end;
    typeRanges[4] = m->subsurfaceNodes.begin; typeRanges[5] = m->subsurfaceNodes.end;
    typeRanges[6] = m->interiorNodes.begin; typeRanges[7] = m->interiorNodes.end;
}

// evaluator: 0 mean ratio, 1 metric conformity; sampler: 0 uniform,
// 1 analytic shell, 2 texture; measurer follows the sampler (MeshCrew.cpp:93-103)
// unless measurer >= 0 is given.
int orc_set_crew(void* h, int evaluator, int sampler, int measurer, double scaling,
                 const int32_t* tsize, const double* tmin, const double* text, const float* texels)
{
    Mesh* m = (Mesh*)h;
    Crew& c = m->crew;
    c.evaluator = evaluator;
    c.sampler = sampler;
    c.measurer = measurer >= 0 ? measurer : (sampler == SAMP_UNIFORM ? MEAS_METRIC_FREE : MEAS_METRIC_WISE);
    c.scaling = scaling; c.scaling2 = c.scaling * scaling; c.scaling3 = c.scaling2 * scaling;
    samplerRotation(c);
    if(sampler == SAMP_TEXTURE)
    {
        for(int d = 0; d < 3; ++d) c.tsize[d] = tsize[d];
        c.tmin = dvec3(tmin[0], tmin[1], tmin[2]);
        c.text = dvec3(text[0], text[1], text[2]);
        size_t n = 6 * (size_t)tsize[0] * tsize[1] * tsize[2];
        c.texels.assign(texels, texels + n);
    }
    return 0;
}

// AbstractSampler::setAspectRatio (AbstractSampler.cpp:56-59).
void orc_set_aspect(void* h, double aspect) { ((Mesh*)h)->crew.aspect = aspect; }

// out: [min, invSum, n], hist[20]
void orc_evaluate(void* h, int chunks, double* out, int32_t* hist)
{
    Histogram hh = evaluateMesh(*(Mesh*)h, chunks);
    out[0] = hh.minQ; out[1] = hh.invSum; out[2] = (double)hh.n;
    for(int i = 0; i < 20; ++i) hist[i] = hh.buckets[i];
}

double orc_tet_quality(void* h, const double* p)
{
    dvec3 vp[4];
    for(int k = 0; k < 4; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return tetQuality(((Mesh*)h)->crew, vp);
}

double orc_hex_quality(void* h, const double* p)
{
    dvec3 vp[8];
    for(int k = 0; k < 8; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return hexQuality(((Mesh*)h)->crew, vp);
}

double orc_pri_quality(void* h, const double* p)
{
    dvec3 vp[6];
    for(int k = 0; k < 6; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return priQuality(((Mesh*)h)->crew, vp);
}

double orc_pri_volume(void* h, const double* p)
{
    dvec3 vp[6];
    for(int k = 0; k < 6; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return priVolume(((Mesh*)h)->crew, vp);
}

// Per-element qualities, tets, then prisms, then hexes (the order of
// evaluateMeshQualitySerial).
void orc_element_qualities(void* h, double* out)
{
    Mesh* m = (Mesh*)h;
    size_t i = 0;
    for(size_t e = 0; e < m->tets.size(); ++e) out[i++] = tetQualityMesh(*m, (uint32_t)e);
    for(size_t e = 0; e < m->pris.size(); ++e) out[i++] = priQualityMesh(*m, (uint32_t)e);
    for(size_t e = 0; e < m->hexs.size(); ++e) out[i++] = hexQualityMesh(*m, (uint32_t)e);
}

// The project libm on the host (op 2 pow23, 3 cbrt, 4 exp, 6 cos, 7 log,
// 8 pow(a, b)), for the device probe test.
void orc_gm_libm2(int op, const double* a, const double* b, double* out, int64_t n)
{
    for(int64_t i = 0; i < n; ++i)
        out[i] = op == 2 ? gm_pow23(a[i]) : op == 3 ? gm_cbrt_nonneg(a[i]) : op == 4 ? gm_exp(a[i])
               : op == 6 ? gm_cos(a[i]) : op == 7 ? gm_log(a[i]) : gm_pow(a[i], b[i]);
}

void orc_gm_libm(int op, const double* in, double* out, int64_t n) { orc_gm_libm2(op, in, in, out, n); }

double orc_patch_quality(void* h, uint32_t vId) { return patchQuality(*(Mesh*)h, vId, nullptr); }

double orc_local_size(void* h, uint32_t vId) { return computeLocalElementSize(*(Mesh*)h, vId); }

void orc_metric_at(void* h, const double* p, double* out9)
{
    dmat3 M = metricAt(((Mesh*)h)->crew, dvec3(p[0], p[1], p[2]));
    for(int j = 0; j < 3; ++j) for(int i = 0; i < 3; ++i) out9[3*j+i] = M[j][i];
}

double orc_riemannian_distance(void* h, const double* a, const double* b)
{
    const Crew& c = ((Mesh*)h)->crew;
    dvec3 A(a[0], a[1], a[2]), B(b[0], b[1], b[2]);
    if(c.measurer == MEAS_METRIC_WISE) return riemannianDistanceWise(c, A, B);
    return length(B - A) * c.scaling;
}

double orc_tet_volume(void* h, const double* p)
{
    dvec3 vp[4];
    for(int k = 0; k < 4; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return tetVolume(((Mesh*)h)->crew, vp);
}

double orc_hex_volume(void* h, const double* p)
{
    dvec3 vp[8];
    for(int k = 0; k < 8; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    return hexVolume(((Mesh*)h)->crew, vp);
}

// Runs one vertex-wise smoother on vertex vId and records every patch
// evaluation: out[4 i .. 4 i + 3] = (x, y, z, quality) of evaluation i, at
// most cap entries; returns the number of evaluations.
int64_t orc_trace_vertex(void* h, int smoother, double spawnCoeff, uint32_t vId, double* out, int64_t cap)
{
    Mesh* m = (Mesh*)h;
    std::vector<double> trace;
    g_trace = &trace;
    switch(smoother)
    {
    case 0: nelderMeadVertex(*m, vId, nullptr); break;
    case 1: gradientDescentVertex(*m, vId, nullptr); break;
    case 2: spawnSearchVertex(*m, vId, spawnCoeff, nullptr); break;
    }
    g_trace = nullptr;
    const int64_t n = (int64_t)trace.size() / 4;
    std::memcpy(out, trace.data(), (size_t)std::min(n, cap) * 4 * sizeof(double));
    return n;
}

// GETMe element step on free-standing vertices (no snapping): kind 0 = tet
// (4 vertices), 1 = hex (8), 2 = prism (6); vpp = transformed positions,
// *weight = q'/(1+q).
void orc_getme_element(void* h, int kind, const double* p, double* vppOut, double* weight)
{
    const Crew& c = ((Mesh*)h)->crew;
    const int n = kind == 0 ? 4 : kind == 2 ? 6 : 8;
    dvec3 vp[8], vpp[8];
    for(int k = 0; k < n; ++k) vp[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    if(kind == 0) getmeTetTransform(c, vp, vpp);
    else if(kind == 2) getmePriTransform(c, vp, vpp);
    else getmeHexTransform(c, vp, vpp);
    const double q = kind == 0 ? tetQuality(c, vp) : kind == 2 ? priQuality(c, vp) : hexQuality(c, vp);
    const double qp = kind == 0 ? tetQuality(c, vpp) : kind == 2 ? priQuality(c, vpp) : hexQuality(c, vpp);
    *weight = qp / (1.0 + q);
    for(int k = 0; k < n; ++k) { vppOut[3*k] = vpp[k].x; vppOut[3*k+1] = vpp[k].y; vppOut[3*k+2] = vpp[k].z; }
}

// MetricConformityEvaluator::specifiedMetric of a tetrahedron, column-major.
void orc_specified_metric(void* h, const double* p, double* out9)
{
    dvec3 v[4];
    for(int k = 0; k < 4; ++k) v[k] = dvec3(p[3*k], p[3*k+1], p[3*k+2]);
    dmat3 M = specifiedMetric(((Mesh*)h)->crew, v[0], v[1], v[2], v[3]);
    for(int j = 0; j < 3; ++j) for(int i = 0; i < 3; ++i) out9[3*j+i] = M[j][i];
}

// Smooths one vertex-wise group list (tests of single-vertex behaviour).
void orc_smooth_vertices(void* h, int smoother, double spawnCoeff, const uint32_t* vids, int64_t n, int64_t* evalCounts)
{
    Mesh* m = (Mesh*)h;
    for(int64_t i = 0; i < n; ++i)
    {
        int64_t c = 0;
        SmoothParams sp; sp.smoother = smoother; sp.spawnMoveCoeff = spawnCoeff;
        std::vector<int64_t> tmp;
        int64_t* cnt = &c;
        int64_t* dummy = cnt;
        (void)dummy;
        switch(smoother)
        {
        case 0: nelderMeadVertex(*m, vids[i], cnt); break;
        case 1: gradientDescentVertex(*m, vids[i], cnt); break;
        case 2: spawnSearchVertex(*m, vids[i], spawnCoeff, cnt); break;
        }
        if(evalCounts) evalCounts[i] = c;
    }
}

// Full smoothing run (AbstractSmoother::smoothMesh with topology operations
// off; AbstractSmoother.cpp:103-211).
//   smoother: 0 NM, 1 GD, 2 spawn search, 3 GETMe.
//   passes: Schedule::relocationPassCount.  With autopilot == 0 exactly
//     `passes` sweeps run.  With autopilot != 0 the sweep after pass record
//     k >= 1 runs only while
//       mean_k - mean_{k-1} > minQualThreshold  or  min_k - min_{k-1} > qualMeanThreshold
//     (AbstractSmoother.cpp:186-193: the harmonic-mean gain is compared with
//     minQualThreshold and the minimum gain with qualMeanThreshold), and after
//     record 0 while 0 < passes (:194-197); the reference has no other bound,
//     here `passes` also caps the sweeps.
//   order: 0 = serialGroup order on one thread (smoothMeshSerial),
//          1 = colour-group-major, `threads` workers per group (smoothMeshThread).
//   evalChunks: 1 = serial histogram sums, k > 1 = k-way chunked sums.
// stats: (passes+1) x [min, harmonicMean, invSum]; hist: (passes+1) x 20
// (records past the last one run are left as they are); evalCounts:
// per-vertex patch evaluation counts (may be null); *passesRun = sweeps run.
int orc_smooth_schedule(void* h, int smoother, int passes, int autopilot, double minQualThreshold,
                        double qualMeanThreshold, int threads, int order, int evalChunks, double spawnCoeff,
                        double* stats, int32_t* hist, int64_t* evalCounts, int* passesRun)
{
    Mesh& m = *(Mesh*)h;
    size_t V = m.verts.size();
    std::vector<int64_t> counts(V, 0);
    SmoothParams sp;
    sp.smoother = smoother;
    sp.spawnMoveCoeff = spawnCoeff;
    sp.evalCount = &counts;

    bool hasBoundaryTypes = (m.boundaryNodes.end > m.boundaryNodes.begin) ||
                            (m.subsurfaceNodes.end > m.subsurfaceNodes.begin);
    double lastMin = 0.0, lastMean = 0.0;
    bool keepGoing = true;
    auto record = [&](int pass) {
        Histogram hh = evaluateMesh(m, evalChunks);
        if(stats) { stats[3*pass] = hh.minQ; stats[3*pass+1] = hh.harmonicMean(); stats[3*pass+2] = hh.invSum; }
        if(hist) for(int i = 0; i < 20; ++i) hist[20*pass+i] = hh.buckets[i];
        if(autopilot && pass > 0)
        {
            const double minGain = hh.minQ - lastMin;
            const double geomGain = hh.harmonicMean() - lastMean;
            keepGoing = (geomGain > minQualThreshold) || (minGain > qualMeanThreshold);
        }
        lastMin = hh.minQ;
        lastMean = hh.harmonicMean();
    };

    GetmeOut gout;
    int swept = 0;
    for(int pass = 0; pass <= passes; ++pass)
    {
        record(pass);
        if(pass == passes || !keepGoing) break;
        ++swept;
        if(smoother == 3)
        {
            gout.kind[TET].resize(m.tets.size());
            gout.kind[PRI].resize(m.pris.size());
            gout.kind[HEX].resize(m.hexs.size());
            parallelFor(m.tets.size(), threads, [&](size_t e) { getmeTet(m, (uint32_t)e, gout.kind[TET][e]); });
            parallelFor(m.pris.size(), threads, [&](size_t e) { getmePri(m, (uint32_t)e, gout.kind[PRI][e]); });
            parallelFor(m.hexs.size(), threads, [&](size_t e) { getmeHex(m, (uint32_t)e, gout.kind[HEX][e]); });
            if(order == 0 || hasBoundaryTypes)
            {
                for(uint32_t v : m.serialGroup) getmeUpdateVertex(m, v, gout, &counts[v]);
            }
            else
            {
                for(const ParallelGroup& g : m.parallelGroups)
                    parallelFor(g.undispatchedNodes.size(), threads, [&](size_t i) {
                        uint32_t v = g.undispatchedNodes[i];
                        getmeUpdateVertex(m, v, gout, &counts[v]);
                    });
            }
        }
        else if(order == 0)
        {
            for(uint32_t v : m.serialGroup) smoothVertex(m, v, sp);
        }
        else
        {
            for(const ParallelGroup& g : m.parallelGroups)
                parallelFor(g.undispatchedNodes.size(), threads, [&](size_t i) { smoothVertex(m, g.undispatchedNodes[i], sp); });
        }
    }
    if(evalCounts) for(size_t v = 0; v < V; ++v) evalCounts[v] = counts[v];
    if(passesRun) *passesRun = swept;
    return 0;
}

// Exactly `passes` sweeps (autopilot off).
int orc_smooth(void* h, int smoother, int passes, int threads, int order, int evalChunks,
               double spawnCoeff, double* stats, int32_t* hist, int64_t* evalCounts)
{
    return orc_smooth_schedule(h, smoother, passes, 0, 0.0, 0.0, threads, order, evalChunks, spawnCoeff, stats,
                               hist, evalCounts, nullptr);
}

} // extern "C"
