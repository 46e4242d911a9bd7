// binding_test.cpp -- include/gpumesh_binding.hpp compiled against
// stand-ins of the reference's Mesh, NodeGroups and QualityHistogram that
// carry the members the binding uses, and driven the way the reference's
// smoothing loop drives it (AbstractSmoother::evaluateMeshQuality,
// Smoothers/AbstractSmoother.cpp:103-211, around one sweep per pass).
//
//   binding_test host <dir>     host topology only: writes the handle's
//                               permutation, elements and groups back
//   binding_test restructure <dir> <K>
//                               host topology only: BATR on the handle, the
//                               mesh handed back as before compileTopology
//   binding_test global <dir> <K> <smoother> <passes>
//                               on the GPU: BATR on the handle, the mesh
//                               renumbered as the reference's compileTopology
//                               would, `passes` FP64 sweeps, positions fetched
//   binding_test device <dir> <smoother> <passes> <precision>
//                               the whole run on the GPU: one evaluation
//                               before the passes, a record per pass,
//                               positions fetched into the mesh at the end
// Inputs (raw arrays written by tests/test_binding.py): pos.bin (f64),
// fixed.bin (i32), tets.bin, hexs.bin (u32), goff.bin (i64), gnodes.bin (u32);
// for sliding vertices also cons.bin (i32, a row of ctab.bin or -1) and
// ctab.bin (7 f64 a row: kind 1 plane through a with normal b, 2 segment
// a-b), which become stand-in PlaneConstraint / SegmentConstraint objects
// described to the binding by their geometry.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gpumesh_binding.hpp"

namespace ref
{
struct dvec3 { double x, y, z; };
struct MeshVert { dvec3 p; };
struct AbstractConstraint
{
    explicit AbstractConstraint(int id) : _id(id) {}
    virtual ~AbstractConstraint() = default;
    int id() const { return _id; }
    bool isFixed() const { return _id < 0; }          // < 0 fixed, 0 free (AbstractConstraint.h:65-73)
    bool isConstrained() const { return _id != 0; }
    int _id;
};
struct PlaneConstraint : AbstractConstraint
{
    PlaneConstraint(int id, dvec3 p, dvec3 n) : AbstractConstraint(id), p(p), n(n) {}
    dvec3 p, n;
};
struct SegmentConstraint : AbstractConstraint
{
    SegmentConstraint(int id, dvec3 a, dvec3 b) : AbstractConstraint(id), a(a), b(b) {}
    dvec3 a, b;
};
struct MeshTopo { const AbstractConstraint* snapToBoundary; };
struct MeshTet { uint32_t v[4]; };
struct MeshHex { uint32_t v[8]; };
struct MeshPri { uint32_t v[6]; };
struct ParallelGroup { std::vector<uint32_t> undispatchedNodes; };
struct NodeGroups
{
    std::vector<ParallelGroup> groups;
    const std::vector<ParallelGroup>& parallelGroups() const { return groups; }
};
struct Mesh
{
    std::vector<MeshVert> verts;
    std::vector<MeshTopo> topos;
    std::vector<MeshTet> tets;
    std::vector<MeshPri> pris;
    std::vector<MeshHex> hexs;
    NodeGroups groups;
    const NodeGroups& nodeGroups() const { return groups; }
};
struct QualityHistogram
{
    std::vector<int> buckets = std::vector<int>(20, 0);
    size_t samples = 0;
    double minimum = 1.0, invSum = 0.0;
    void setMinimumQuality(double m) { minimum = m; }
    void setInvQualitySum(double s) { invSum = s; }
    void setSampleCount(size_t n) { samples = n; }
    void setBucket(size_t i, int c) { buckets[i] = c; }
};
} // namespace ref

template<typename X> static std::vector<X> load(const std::string& dir, const char* name)
{
    std::vector<X> v;
    FILE* f = std::fopen((dir + "/" + name).c_str(), "rb");
    if(!f) { std::fprintf(stderr, "missing %s\n", name); std::exit(2); }
    X x;
    while(std::fread(&x, sizeof(X), 1, f) == 1) v.push_back(x);
    std::fclose(f);
    return v;
}

template<typename X> static void save(const std::string& dir, const char* name, const X* p, size_t n)
{
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    std::fwrite(p, sizeof(X), n, f);
    std::fclose(f);
}

static const ref::AbstractConstraint FIXED{-1}, FREE{0};
static std::vector<std::unique_ptr<ref::AbstractConstraint>> g_sliding;

// What an integrator writes for its boundary: the geometry of the plane and
// segment constraints, nothing for any other kind.
struct Describe
{
    bool operator()(const ref::AbstractConstraint& c, gm_constraint& out) const
    {
        if(const auto* p = dynamic_cast<const ref::PlaneConstraint*>(&c))
        {
            out.kind = GM_CONSTRAINT_PLANE;
            out.a[0] = p->p.x; out.a[1] = p->p.y; out.a[2] = p->p.z;
            out.b[0] = p->n.x; out.b[1] = p->n.y; out.b[2] = p->n.z;
            return true;
        }
        if(const auto* e = dynamic_cast<const ref::SegmentConstraint*>(&c))
        {
            out.kind = GM_CONSTRAINT_SEGMENT;
            out.a[0] = e->a.x; out.a[1] = e->a.y; out.a[2] = e->a.z;
            out.b[0] = e->b.x; out.b[1] = e->b.y; out.b[2] = e->b.z;
            return true;
        }
        return false;
    }
};

static bool exists(const std::string& dir, const char* name)
{
    FILE* f = std::fopen((dir + "/" + name).c_str(), "rb");
    if(f) std::fclose(f);
    return f != nullptr;
}

static ref::Mesh loadMesh(const std::string& dir)
{
    ref::Mesh m;
    const std::vector<double> pos = load<double>(dir, "pos.bin");
    const std::vector<int32_t> fixed = load<int32_t>(dir, "fixed.bin");
    const std::vector<uint32_t> tets = load<uint32_t>(dir, "tets.bin");
    const std::vector<uint32_t> hexs = load<uint32_t>(dir, "hexs.bin");
    const std::vector<int64_t> goff = load<int64_t>(dir, "goff.bin");
    const std::vector<uint32_t> gnodes = load<uint32_t>(dir, "gnodes.bin");
    std::vector<int32_t> cons;
    if(exists(dir, "cons.bin"))
    {
        cons = load<int32_t>(dir, "cons.bin");
        const std::vector<double> t = load<double>(dir, "ctab.bin");
        for(size_t k = 0; 7 * k < t.size(); ++k)
        {
            const double* r = &t[7 * k];
            const ref::dvec3 a{r[1], r[2], r[3]}, b{r[4], r[5], r[6]};
            if((int)r[0] == 1) g_sliding.emplace_back(new ref::PlaneConstraint((int)k + 1, a, b));
            else g_sliding.emplace_back(new ref::SegmentConstraint((int)k + 1, a, b));
        }
    }
    m.verts.resize(fixed.size());
    m.topos.resize(fixed.size());
    for(size_t v = 0; v < fixed.size(); ++v)
    {
        m.verts[v].p = ref::dvec3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]};
        m.topos[v].snapToBoundary = fixed[v] ? &FIXED : !cons.empty() && cons[v] >= 0 ? g_sliding[cons[v]].get() : &FREE;
    }
    m.tets.resize(tets.size() / 4);
    for(size_t e = 0; e < m.tets.size(); ++e) for(int k = 0; k < 4; ++k) m.tets[e].v[k] = tets[4 * e + k];
    m.hexs.resize(hexs.size() / 8);
    for(size_t e = 0; e < m.hexs.size(); ++e) for(int k = 0; k < 8; ++k) m.hexs[e].v[k] = hexs[8 * e + k];
    for(size_t g = 0; g + 1 < goff.size(); ++g)
        m.groups.groups.push_back(ref::ParallelGroup{std::vector<uint32_t>(gnodes.begin() + goff[g], gnodes.begin() + goff[g + 1])});
    return m;
}

int main(int argc, char** argv)
{
    if(argc < 3) return 2;
    const std::string mode = argv[1], dir = argv[2];
    ref::Mesh mesh = loadMesh(dir);
    try
    {
        const bool sliding = !g_sliding.empty();
        if(mode == "host")
        {
            if(sliding)
            {
                bool refused = false;
                try { gpumesh_binding::Session<ref::Mesh> t(mesh, false); }
                catch(const std::exception& e) { refused = true; std::printf("undescribed refused: %s\n", e.what()); }
                if(!refused) { std::fprintf(stderr, "sliding vertices accepted without a description\n"); return 3; }
            }
            gpumesh_binding::Session<ref::Mesh> s(mesh, Describe(), false);
            const size_t V = mesh.verts.size();
            std::vector<int64_t> perm(V);
            gpumesh_binding::check(gm_mesh_permutation(s.handle(), perm.data()), "gm_mesh_permutation");
            std::vector<uint32_t> tets(4 * mesh.tets.size() + 1), hexs(8 * mesh.hexs.size() + 1);
            gpumesh_binding::check(gm_mesh_elements(s.handle(), tets.data(), hexs.data()), "gm_mesh_elements");
            std::vector<int64_t> goff(s.groups() + 1);
            std::vector<uint32_t> gnodes(V + 1);
            gpumesh_binding::check(gm_mesh_groups(s.handle(), goff.data(), gnodes.data()), "gm_mesh_groups");
            save(dir, "out_perm.bin", perm.data(), V);
            save(dir, "out_tets.bin", tets.data(), 4 * mesh.tets.size());
            save(dir, "out_hexs.bin", hexs.data(), 8 * mesh.hexs.size());
            save(dir, "out_goff.bin", goff.data(), goff.size());
            save(dir, "out_gnodes.bin", gnodes.data(), (size_t)goff.back());
            // A colouring that puts two element-mates in one group is refused:
            // move a free vertex into the group of a free element-mate.
            uint32_t a = 0, b = 0;
            bool found = false;
            for(const auto& t : mesh.tets)
            {
                for(int i = 0; i < 4 && !found; ++i)
                    for(int j = i + 1; j < 4 && !found; ++j)
                        if(!mesh.topos[t.v[i]].snapToBoundary->isFixed() && !mesh.topos[t.v[j]].snapToBoundary->isFixed())
                        {
                            a = t.v[i];
                            b = t.v[j];
                            found = true;
                        }
                if(found) break;
            }
            if(found)
            {
                ref::Mesh bad = mesh;
                ref::ParallelGroup* ga = nullptr;
                for(auto& g : bad.groups.groups)
                    for(uint32_t x : g.undispatchedNodes)
                        if(x == a) ga = &g;
                for(auto& g : bad.groups.groups)
                    for(auto it = g.undispatchedNodes.begin(); it != g.undispatchedNodes.end(); ++it)
                        if(*it == b) { g.undispatchedNodes.erase(it); break; }
                ga->undispatchedNodes.push_back(b);
                bool refused = false;
                try { gpumesh_binding::Session<ref::Mesh> t(bad, Describe(), false); }
                catch(const std::exception& e) { refused = true; std::printf("refused: %s\n", e.what()); }
                if(!refused) { std::fprintf(stderr, "shared-element group accepted\n"); return 3; }
            }
            std::printf("host ok\n");
            return 0;
        }
        if(mode == "restructure" && argc >= 4)
        {
            gpumesh_binding::Session<ref::Mesh> s(mesh, Describe(), false);
            s.setCrew(GM_MEAN_RATIO, GM_UNIFORM, std::atof(argv[3]));
            const std::vector<int64_t> counts = s.restructure(
                mesh, 5, 5, 0.5, 1.5,
                [](bool fixed, double, double, double) { return ref::MeshTopo{fixed ? &FIXED : &FREE}; });
            const size_t V = mesh.verts.size();
            std::vector<double> pos(3 * V);
            std::vector<int32_t> fixed(V);
            for(size_t v = 0; v < V; ++v)
            {
                pos[3 * v] = mesh.verts[v].p.x;
                pos[3 * v + 1] = mesh.verts[v].p.y;
                pos[3 * v + 2] = mesh.verts[v].p.z;
                fixed[v] = mesh.topos[v].snapToBoundary->isFixed() ? 1 : 0;
            }
            std::vector<uint32_t> tets;
            for(const auto& t : mesh.tets) tets.insert(tets.end(), t.v, t.v + 4);
            save(dir, "out_counts.bin", counts.data(), counts.size());
            save(dir, "out_pos.bin", pos.data(), pos.size());
            save(dir, "out_fixed.bin", fixed.data(), fixed.size());
            save(dir, "out_tets.bin", tets.data(), tets.size());
            std::printf("restructure ok\n");
            return 0;
        }
        if(mode == "global" && argc >= 6)
        {
            const int smoother = std::atoi(argv[4]), passes = std::atoi(argv[5]);
            {
                gpumesh_binding::Session<ref::Mesh> s(mesh, Describe(), true);
                s.setCrew(GM_MEAN_RATIO, GM_UNIFORM, std::atof(argv[3]));
                s.upload(mesh);
                const std::vector<int64_t> counts = s.restructure(
                    mesh, 5, 5, 0.5, 1.5,
                    [](bool fixed, double, double, double) { return ref::MeshTopo{fixed ? &FIXED : &FREE}; });
                save(dir, "out_counts.bin", counts.data(), counts.size());
                // Mesh::compileTopology's renumbering, which the handle has.
                const size_t V = mesh.verts.size();
                std::vector<int64_t> perm(V);
                gpumesh_binding::check(gm_mesh_permutation(s.handle(), perm.data()), "gm_mesh_permutation");
                ref::Mesh compiled = mesh;
                for(size_t v = 0; v < V; ++v)
                {
                    compiled.verts[(size_t)perm[v]] = mesh.verts[v];
                    compiled.topos[(size_t)perm[v]] = mesh.topos[v];
                }
                for(auto& t : compiled.tets)
                    for(uint32_t& c : t.v) c = (uint32_t)perm[c];
                mesh = compiled;
                for(int p = 0; p < passes; ++p) s.sweep(smoother, GM_FP64, 1.0 / 20.0);
                s.fetch(mesh);
            }
            std::vector<double> pos(3 * mesh.verts.size());
            for(size_t v = 0; v < mesh.verts.size(); ++v)
            {
                pos[3 * v] = mesh.verts[v].p.x;
                pos[3 * v + 1] = mesh.verts[v].p.y;
                pos[3 * v + 2] = mesh.verts[v].p.z;
            }
            std::vector<uint32_t> tets;
            for(const auto& t : mesh.tets) tets.insert(tets.end(), t.v, t.v + 4);
            save(dir, "out_pos.bin", pos.data(), pos.size());
            save(dir, "out_tets.bin", tets.data(), tets.size());
            std::printf("global ok\n");
            return 0;
        }
        if(mode == "device" && argc >= 6)
        {
            const int smoother = std::atoi(argv[3]), passes = std::atoi(argv[4]), precision = std::atoi(argv[5]);
            std::vector<double> stats;
            std::vector<int32_t> hist;
            {
                gpumesh_binding::Session<ref::Mesh> s(mesh, Describe(), true);
                s.setCrew(GM_MEAN_RATIO, GM_UNIFORM, 1.0);
                s.upload(mesh);
                // The reference's loop: an initial evaluation, then a record
                // per pass around one sweep each (topology operations off).
                ref::QualityHistogram h;
                auto* active = gpumesh_binding::Session<ref::Mesh>::of(&mesh);
                if(active != &s) { std::fprintf(stderr, "session not found\n"); return 4; }
                active->evaluate(h);
                for(int p = 0;; ++p)
                {
                    active->evaluate(h);
                    stats.push_back(h.minimum);
                    stats.push_back(h.invSum);
                    stats.push_back((double)h.samples);
                    hist.insert(hist.end(), h.buckets.begin(), h.buckets.end());
                    if(p >= passes) break;
                    s.sweep(smoother, precision, 1.0 / 20.0);
                }
                s.fetch(mesh);
            }
            std::vector<double> pos(3 * mesh.verts.size());
            for(size_t v = 0; v < mesh.verts.size(); ++v)
            {
                pos[3 * v] = mesh.verts[v].p.x;
                pos[3 * v + 1] = mesh.verts[v].p.y;
                pos[3 * v + 2] = mesh.verts[v].p.z;
            }
            save(dir, "out_pos.bin", pos.data(), pos.size());
            save(dir, "out_stats.bin", stats.data(), stats.size());
            save(dir, "out_hist.bin", hist.data(), hist.size());
            std::printf("device ok\n");
            return 0;
        }
    }
    catch(const std::exception& e)
    {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 2;
}
