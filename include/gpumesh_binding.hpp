/* gpumesh_binding.hpp -- the reference-side binding of the C ABI
 * (gpumesh.h), header-only C++.
 *
 * It is written against the reference's own types as template parameters,
 * using only the members the reference declares for them, so it compiles
 * into the reference without a build dependency and into the tests against
 * stand-ins with the same members:
 *   Mesh:       verts[v].p.{x,y,z} (DataStructures/Mesh.h:25-37),
 *               topos[v].snapToBoundary->isFixed() / ->isConstrained() / ->id()
 *               (Mesh.h:256-267, Boundaries/Constraints/AbstractConstraint.h:24,65-73),
 *               tets[e].v[4], hexs[e].v[8] (Mesh.h:66-97, 199-234),
 *               nodeGroups().parallelGroups()[g].undispatchedNodes
 *               (DataStructures/NodeGroups.h:71-93);
 *   Histogram:  setMinimumQuality, setInvQualitySum, setSampleCount,
 *               setBucket (DataStructures/QualityHistogram.h:20-34).
 *
 * The device copy keeps the Mesh's own numbering and its NodeGroups colour
 * groups (gm_mesh_create_desc with groups), so positions go back to the
 * vertices they came from without any permutation.
 *
 * Vertices that slide on a face or an edge need their constraint's
 * geometry, which PlaneConstraint and SegmentConstraint keep private
 * (FaceConstraint.h:37-45, EdgeConstraint.h:31-40).  The caller passes a
 * `describe(const Constraint&, gm_constraint&) -> bool` function that knows
 * it (the Boundary that made the constraints does: BoxBoundary.cpp:7-36);
 * constraints are shared by id(), and one that describe() declines (a curved
 * face, say) makes the constructor throw, as does any sliding vertex when no
 * describe() is given.
 */
#ifndef GPUMESH_BINDING_HPP
#define GPUMESH_BINDING_HPP

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpumesh.h"

namespace gpumesh_binding
{

inline void check(int rc, const char* where)
{
    if(rc != GM_OK) throw std::runtime_error(std::string(where) + ": " + gm_last_error());
}

// Positions of a reference Mesh as 3 doubles per vertex.
template<class MeshT> std::vector<double> positionsOf(const MeshT& mesh)
{
    const size_t V = mesh.verts.size();
    std::vector<double> pos(3 * V);
    for(size_t v = 0; v < V; ++v)
    {
        pos[3 * v] = mesh.verts[v].p.x;
        pos[3 * v + 1] = mesh.verts[v].p.y;
        pos[3 * v + 2] = mesh.verts[v].p.z;
    }
    return pos;
}

// A describe() that knows no constraint: every sliding vertex is refused.
struct NoSlidingConstraints
{
    template<class C> bool operator()(const C&, gm_constraint&) const { return false; }
};

// The flat arrays of a reference Mesh, in its own numbering.
template<class MeshT> struct MeshArrays
{
    std::vector<double> pos;
    std::vector<int32_t> fixed;
    std::vector<int32_t> cons;                // constraint table index per vertex, -1 for none
    std::vector<gm_constraint> table;
    std::vector<uint32_t> tets, hexs, pris;
    std::vector<int64_t> groupOff;
    std::vector<uint32_t> groupNodes;

    template<class Describe> MeshArrays(const MeshT& mesh, const Describe& describe) : pos(positionsOf(mesh))
    {
        const size_t V = mesh.verts.size();
        fixed.resize(V);
        cons.assign(V, -1);
        std::map<int, int32_t> byId;
        for(size_t v = 0; v < V; ++v)
        {
            const auto& c = *mesh.topos[v].snapToBoundary;
            fixed[v] = c.isFixed() ? 1 : 0;
            if(!c.isConstrained() || c.isFixed()) continue;
            auto it = byId.find(c.id());
            if(it == byId.end())
            {
                gm_constraint g{};
                if(!describe(c, g))
                    throw std::runtime_error("vertex " + std::to_string(v) + " slides on constraint " +
                                             std::to_string(c.id()) + ", which has no plane or segment "
                                             "description: smooth this mesh with the Thread implementation");
                it = byId.emplace(c.id(), (int32_t)table.size()).first;
                table.push_back(g);
            }
            cons[v] = it->second;
        }
        tets.reserve(4 * mesh.tets.size());
        for(const auto& t : mesh.tets) tets.insert(tets.end(), t.v, t.v + 4);
        hexs.reserve(8 * mesh.hexs.size());
        for(const auto& h : mesh.hexs) hexs.insert(hexs.end(), h.v, h.v + 8);
        pris.reserve(6 * mesh.pris.size());
        for(const auto& p : mesh.pris) pris.insert(pris.end(), p.v, p.v + 6);
        groupOff.push_back(0);
        for(const auto& g : mesh.nodeGroups().parallelGroups())
        {
            groupNodes.insert(groupNodes.end(), g.undispatchedNodes.begin(), g.undispatchedNodes.end());
            groupOff.push_back((int64_t)groupNodes.size());
        }
    }
};

// A device-resident copy of one Mesh for the span of a smoothing call (the
// reference's CUDA path likewise holds one mesh at a time in its device
// globals).  device = false builds the host topology only.
template<class MeshT> class Session
{
public:
    explicit Session(const MeshT& mesh, bool device = true) : Session(mesh, NoSlidingConstraints(), device) {}

    template<class Describe> Session(const MeshT& mesh, const Describe& describe, bool device = true) : _mesh(&mesh)
    {
        MeshArrays<MeshT> a(mesh, describe);
        gm_mesh_desc d{};
        d.nverts = (int64_t)mesh.verts.size();
        d.pos = a.pos.data();
        d.fixed = a.fixed.data();
        d.ntets = (int64_t)mesh.tets.size();
        d.tets = a.tets.data();
        d.nhexs = (int64_t)mesh.hexs.size();
        d.hexs = a.hexs.data();
        d.npris = (int64_t)mesh.pris.size();
        d.pris = a.pris.data();
        d.constraint = a.table.empty() ? nullptr : a.cons.data();
        d.nconstraints = (int64_t)a.table.size();
        d.constraints = a.table.empty() ? nullptr : a.table.data();
        d.ngroups = (int64_t)a.groupOff.size() - 1;
        d.group_off = a.groupOff.data();
        d.group_nodes = a.groupNodes.data();
        d.device_copy = device ? 1 : 0;
        check(gm_mesh_create_desc(&d, &_h), "gm_mesh_create_desc");
        int64_t info[8];
        check(gm_mesh_info(_h, info), "gm_mesh_info");
        _groups = info[4];
        previous() = active();
        active() = this;
    }
    ~Session()
    {
        if(active() == this) active() = previous();
        gm_mesh_destroy(_h);
    }
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;

    gm_mesh* handle() const { return _h; }
    const MeshT* mesh() const { return _mesh; }
    int64_t groups() const { return _groups; }

    // The session of `mesh`, if one is open (the evaluator uses it to read
    // the device positions during a smoothing call).
    static Session* of(const MeshT* mesh) { return active() && active()->_mesh == mesh ? active() : nullptr; }

    // Crew (MeshCrew's evaluator and sampler, gm_mesh_set_crew).
    void setCrew(int evaluator, int sampler, double scaling)
    {
        check(gm_mesh_set_crew(_h, evaluator, sampler, scaling, nullptr, nullptr, nullptr, nullptr),
              "gm_mesh_set_crew");
    }

    // mesh.verts -> device (updateCudaVerts).
    void upload(const MeshT& mesh)
    {
        const std::vector<double> pos = positionsOf(mesh);
        check(gm_mesh_set_positions(_h, pos.data()), "gm_mesh_set_positions");
    }

    // Device -> mesh.verts (fetchCudaVerts), every position to its own vertex.
    void fetch(MeshT& mesh) const
    {
        std::vector<double> pos(3 * mesh.verts.size());
        check(gm_mesh_get_positions(_h, pos.data()), "gm_mesh_get_positions");
        for(size_t v = 0; v < mesh.verts.size(); ++v)
        {
            mesh.verts[v].p.x = pos[3 * v];
            mesh.verts[v].p.y = pos[3 * v + 1];
            mesh.verts[v].p.z = pos[3 * v + 2];
        }
    }

    // One smoothing sweep over every colour group, no quality evaluation
    // (the body of one relocation pass).
    void sweep(int smoother, int precision, double spawnCoeff)
    {
        check(gm_mesh_smooth_groups(_h, smoother, 0, _groups, precision, GM_KERNEL_WAVE, spawnCoeff, nullptr),
              "gm_mesh_smooth_groups");
    }

    // BATR topology operations (BatrTopologist::restructureMesh,
    // Topologists/BatrTopologist.cpp:76-128) on the handle, which is rebuilt
    // in place (gm_mesh_restructure).  `mesh` receives the result as the
    // reference's restructure leaves it just before its Mesh::compileTopology:
    // positions and tets in that numbering (the handle's permutation maps it
    // to the handle's), and each vertex's topology entry from
    // makeTopo(fixed, x, y, z) -- a MeshTopo holding the vertex's own fixed
    // point, or a free one.  The caller then runs the reference's
    // compileTopology, which renumbers the mesh as the handle already is, so
    // fetch/upload/sweep go on without a new session.  Returns the face swap,
    // edge swap, split, merge and pass counts.
    template<class MakeTopo>
    std::vector<int64_t> restructure(MeshT& mesh, int topoPasses, int refinementSweeps, double minEdge,
                                     double maxEdge, const MakeTopo& makeTopo)
    {
        std::vector<int64_t> counts(5, 0);
        check(gm_mesh_restructure(_h, topoPasses, refinementSweeps, minEdge, maxEdge, counts.data()),
              "gm_mesh_restructure");
        if(counts[4] == 0) return counts;   // prisms or hexahedra, or no tets: left as it was
        int64_t info[8];
        check(gm_mesh_info(_h, info), "gm_mesh_info");
        const size_t V = (size_t)info[0], T = (size_t)info[1];
        const int64_t nfixed = info[0] - info[3];   // fixed vertices come first
        _groups = info[4];
        std::vector<int64_t> perm(V);
        check(gm_mesh_permutation(_h, perm.data()), "gm_mesh_permutation");
        std::vector<double> pos(3 * V);
        check(gm_mesh_get_positions(_h, pos.data()), "gm_mesh_get_positions");
        std::vector<uint32_t> tets(4 * T + 1), hexs(1);
        check(gm_mesh_elements(_h, tets.data(), hexs.data()), "gm_mesh_elements");
        std::vector<uint32_t> before(V);
        for(size_t v = 0; v < V; ++v) before[(size_t)perm[v]] = (uint32_t)v;
        mesh.verts.resize(V);
        mesh.topos.resize(V);
        for(size_t v = 0; v < V; ++v)
        {
            const double* p = &pos[3 * (size_t)perm[v]];
            mesh.verts[v].p.x = p[0];
            mesh.verts[v].p.y = p[1];
            mesh.verts[v].p.z = p[2];
            mesh.topos[v] = makeTopo(perm[v] < nfixed, p[0], p[1], p[2]);
        }
        mesh.tets.resize(T);
        for(size_t e = 0; e < T; ++e)
            for(int k = 0; k < 4; ++k) mesh.tets[e].v[k] = before[tets[4 * e + k]];
        return counts;
    }

    // The device positions' quality into a QualityHistogram-like object.
    template<class HistogramT> void evaluate(HistogramT& histogram) const
    {
        double r[3];
        int32_t h[20];
        check(gm_mesh_evaluate(_h, r, h, nullptr), "gm_mesh_evaluate");
        histogram.setMinimumQuality(r[0]);
        histogram.setInvQualitySum(r[1]);
        histogram.setSampleCount((std::size_t)r[2]);
        for(int b = 0; b < 20; ++b) histogram.setBucket((std::size_t)b, h[b]);
    }

private:
    static Session*& active()
    {
        static Session* s = nullptr;
        return s;
    }
    Session*& previous() { return _previous; }

    gm_mesh* _h = nullptr;
    const MeshT* _mesh = nullptr;
    Session* _previous = nullptr;
    int64_t _groups = 0;
};

} // namespace gpumesh_binding

#endif // GPUMESH_BINDING_HPP
