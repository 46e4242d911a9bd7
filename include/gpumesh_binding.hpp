/* gpumesh_binding.hpp -- the reference-side binding of the C ABI
 * (gpumesh.h), header-only C++.
 *
 * It is written against the reference's own types as template parameters,
 * using only the members the reference declares for them, so it compiles
 * into the reference without a build dependency and into the tests against
 * stand-ins with the same members:
 *   Mesh:       verts[v].p.{x,y,z} (DataStructures/Mesh.h:25-37),
 *               topos[v].snapToBoundary->isFixed() / ->isConstrained()
 *               (Mesh.h:256-267, Boundaries/Constraints/AbstractConstraint.h:65-73),
 *               tets[e].v[4], hexs[e].v[8] (Mesh.h:66-97, 199-234),
 *               nodeGroups().parallelGroups()[g].undispatchedNodes
 *               (DataStructures/NodeGroups.h:71-93);
 *   Histogram:  setMinimumQuality, setInvQualitySum, setSampleCount,
 *               setBucket (DataStructures/QualityHistogram.h:20-34).
 *
 * The device copy keeps the Mesh's own numbering and its NodeGroups colour
 * groups (gm_mesh_create_grouped), so positions go back to the vertices they
 * came from without any permutation.  How AbstractVertexWiseSmoother::
 * smoothMeshCuda and AbstractEvaluator::evaluateMeshQualityCuda use it is
 * shown in INTEGRATION.md.
 */
#ifndef GPUMESH_BINDING_HPP
#define GPUMESH_BINDING_HPP

#include <cstdint>
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

// The flat arrays of a reference Mesh, in its own numbering.
template<class MeshT> struct MeshArrays
{
    std::vector<double> pos;
    std::vector<int32_t> fixed;
    std::vector<uint32_t> tets, hexs;
    std::vector<int64_t> groupOff;
    std::vector<uint32_t> groupNodes;

    explicit MeshArrays(const MeshT& mesh)
    {
        const size_t V = mesh.verts.size();
        pos.resize(3 * V);
        fixed.resize(V);
        for(size_t v = 0; v < V; ++v)
        {
            pos[3 * v] = mesh.verts[v].p.x;
            pos[3 * v + 1] = mesh.verts[v].p.y;
            pos[3 * v + 2] = mesh.verts[v].p.z;
            const auto& c = *mesh.topos[v].snapToBoundary;
            if(c.isConstrained() && !c.isFixed())
                throw std::runtime_error("vertex " + std::to_string(v) + " slides on a boundary constraint: "
                                         "smooth this mesh with the Thread implementation");
            fixed[v] = c.isFixed() ? 1 : 0;
        }
        tets.reserve(4 * mesh.tets.size());
        for(const auto& t : mesh.tets) tets.insert(tets.end(), t.v, t.v + 4);
        hexs.reserve(8 * mesh.hexs.size());
        for(const auto& h : mesh.hexs) hexs.insert(hexs.end(), h.v, h.v + 8);
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
    explicit Session(const MeshT& mesh, bool device = true) : _mesh(&mesh)
    {
        MeshArrays<MeshT> a(mesh);
        check(gm_mesh_create_grouped((int64_t)mesh.verts.size(), a.pos.data(), a.fixed.data(),
                                     (int64_t)mesh.tets.size(), a.tets.data(), (int64_t)mesh.hexs.size(),
                                     a.hexs.data(), (int64_t)a.groupOff.size() - 1, a.groupOff.data(),
                                     a.groupNodes.data(), device ? 1 : 0, &_h),
              "gm_mesh_create_grouped");
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
        MeshArrays<MeshT> a(mesh);
        check(gm_mesh_set_positions(_h, a.pos.data()), "gm_mesh_set_positions");
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
