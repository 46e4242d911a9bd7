// gm_topology.h -- host-side mesh topology and independent node groups.
//
// Produces exactly what the reference computes in Mesh::compileNeighborhoods
// (DataStructures/Mesh.cpp:365-460) and NodeGroups::build
// (DataStructures/NodeGroups.cpp:104-619): patch element lists, neighbour
// vertex lists, the greedy colouring, the (type, colour, patch size) sort and
// the vertex renumbering.  The data structures are flat arrays (CSR) instead
// of per-vertex std::vector/std::set, so a million-vertex mesh builds in a
// fraction of a second, with bit-identical integer output.
#ifndef GM_TOPOLOGY_H
#define GM_TOPOLOGY_H

#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

namespace gmt
{

enum { TYPE_FIXED = 0, TYPE_BOUNDARY = 1, TYPE_SUBSURFACE = 2, TYPE_INTERIOR = 3 };

struct Topology
{
    // Inputs after renumbering (new vertex ids).
    int64_t nverts = 0, ntets = 0, nhexs = 0, npris = 0;
    std::vector<double> pos;          // 3 * nverts
    std::vector<uint8_t> fixed;       // nverts
    std::vector<uint32_t> tets;       // 4 * ntets
    // 8 * (nhexs + npris) "records": the hexahedra, then the prisms
    // (v0 .. v5, then v5 twice), so both kinds share one 8-slot layout.
    std::vector<uint32_t> hexs;

    // Patch element lists: eref packs (element id << 4) | (is_record << 3) |
    // local index; a record id below nhexs is a hexahedron, nhexs + p is
    // prism p.
    std::vector<uint32_t> eoff, eref;
    // Neighbour vertex lists (edge neighbours, in discovery order).
    std::vector<uint32_t> noff, nbr;

    // Node groups.
    std::vector<int64_t> permutation;      // old id -> new id
    int64_t nfixed = 0;                    // number of fixed vertices ([0, nfixed) after renumbering)
    std::vector<int64_t> groupOff;         // ngroups + 1 offsets into groupNodes
    std::vector<uint32_t> groupNodes;      // movable vertices, colour by colour (any ids)
    std::vector<int32_t> groupMaxPatch;    // largest patch per group
    std::vector<uint32_t> serialGroup;     // movable vertices in serial order

    // NodeGroups' vertex types (NodeGroups.cpp:124-241), per vertex:
    // TYPE_FIXED, TYPE_BOUNDARY (slides on a face or edge constraint),
    // TYPE_SUBSURFACE (free, in an element with a boundary vertex) or
    // TYPE_INTERIOR.
    std::vector<uint8_t> vtype;
    // Sliding constraint of each vertex: an index into the mesh's constraint
    // table, -1 for none; empty when no vertex slides.
    std::vector<int32_t> cons;
    // Each group's members ordered by type: group g's boundary, subsurface
    // and interior vertices are groupNodes[runOff[3g + t] .. runOff[3g + t + 1])
    // for t = 0, 1, 2 (runOff[3g] = groupOff[g]); runMaxPatch[3g + t] is the
    // largest patch of run (g, t).
    std::vector<int64_t> runOff;
    std::vector<int32_t> runMaxPatch;

    // The runtime's device-side layout of this topology (its own type),
    // built on the first upload and shared by later meshes of a cached
    // topology.
    mutable std::shared_ptr<const void> deviceLayout;

    int64_t ngroups() const { return (int64_t)groupOff.size() - 1; }
};

// Builds the topology.  `fixed[v]` != 0 marks a fixed vertex; `cons` (may
// be null) gives each vertex's sliding constraint (an index >= 0, or -1);
// every other vertex is free.  Returns false and sets `err` on invalid input.
// npris prisms (6 vertices each, MeshPri order) may follow the hexahedra.
bool build(int64_t nverts, const double* pos, const int32_t* fixed, const int32_t* cons,
           int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
           int64_t npris, const uint32_t* pris, Topology& out, std::string& err);

// The topology in the caller's numbering with the caller's colour groups
// (the reference's NodeGroups::parallelGroups): group g holds
// groupNodes[groupOff[g] .. groupOff[g + 1]).  Checks that every movable
// vertex is in exactly one group, no fixed vertex is in one, and no two
// vertices of a group share an element.  permutation is the identity.
// Each group's members are reordered by type (boundary, subsurface,
// interior), keeping their order within a type.
bool buildGrouped(int64_t nverts, const double* pos, const int32_t* fixed, const int32_t* cons, int64_t ntets,
                  const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int64_t npris, const uint32_t* pris,
                  int64_t ngroups, const int64_t* groupOff, const uint32_t* groupNodes, Topology& out,
                  std::string& err);

// build() through the shared topology cache (gm_topocache.cpp): returns the
// topology with `pos` left empty and the renumbered positions in
// `renumbered`; nullptr (and `err`) on invalid input.  With a zero capacity
// (the default) every call builds.
std::shared_ptr<const Topology> buildShared(int64_t nverts, const double* pos, const int32_t* fixed,
                                            const int32_t* cons, int64_t ntets, const uint32_t* tets, int64_t nhexs,
                                            const uint32_t* hexs, int64_t npris, const uint32_t* pris,
                                            std::vector<double>& renumbered, std::string& err);
void setCacheCapacity(size_t bytes);
size_t cacheCapacity();
// Whether the cache holds topology t.
bool isCached(const Topology* t);
// {hits, misses, entries, bytes}
void cacheStats(int64_t* out4);

} // namespace gmt

#endif // GM_TOPOLOGY_H
