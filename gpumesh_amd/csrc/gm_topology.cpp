// gm_topology.cpp -- see gm_topology.h.
#include "gm_topology.h"

#include <algorithm>
#include <numeric>
#include <string>

namespace gmt
{

// Element edge tables: MeshTet::edges, MeshPri::edges and MeshHex::edges
// (Mesh.cpp:21-28, 70-80, 98-111).
static const int TET_EDGES[6][2] = {{0,1},{0,2},{0,3},{1,2},{1,3},{2,3}};
static const int PRI_EDGES[9][2] = {{0,1},{0,2},{1,2},{0,3},{1,4},{2,5},{3,4},{3,5},{4,5}};
static const int HEX_EDGES[12][2] = {{0,1},{0,3},{1,2},{2,3},{0,4},{1,5},{2,6},{3,7},{4,5},{4,7},{5,6},{6,7}};

static const int NO_GROUP = -2;
static const int UNSET_GROUP = -1;

// Patch element lists and neighbour lists in the input numbering
// (Mesh::compileNeighborhoods and addEdge, Mesh.cpp:365-460).
struct Adjacency
{
    std::vector<uint32_t> eoff, eref;   // CSR patch lists, eref = (id << 4) | (is_record << 3) | local
    std::vector<uint32_t> ncnt, ntmp;   // neighbour count, and 3 eoff[v] slots per vertex
};

// The element records of a mesh: 8 slots per hexahedron, then per prism
// (its 6 vertices, then the last one twice).
static std::vector<uint32_t> makeRecords(int64_t H, const uint32_t* hexs, int64_t P, const uint32_t* pris)
{
    std::vector<uint32_t> rec(8 * (size_t)(H + P));
    if(H) std::copy(hexs, hexs + 8 * H, rec.begin());
    for(int64_t p = 0; p < P; ++p)
        for(int k = 0; k < 8; ++k) rec[8 * (size_t)(H + p) + k] = pris[6 * p + (k < 6 ? k : 5)];
    return rec;
}

// Vertices of record r (hexahedra below H, prisms from H on).
static inline int recordVerts(int64_t H, uint32_t r) { return (int64_t)r < H ? 8 : 6; }

static bool checkInputs(int64_t V, int64_t T, const uint32_t* tets, int64_t H, const uint32_t* hexs, int64_t P,
                        const uint32_t* pris, std::string& err)
{
    if(V < 0 || T < 0 || H < 0 || P < 0) { err = "negative size"; return false; }
    if(V > (int64_t)0xFFFFFFF0) { err = "too many vertices"; return false; }
    if(T >= (int64_t)1 << 28 || H + P >= (int64_t)1 << 28)
        { err = "too many elements (limit 2^28 tetrahedra, 2^28 hexahedra and prisms together)"; return false; }
    for(int64_t i = 0; i < 4 * T; ++i) if((int64_t)tets[i] >= V) { err = "tet vertex id out of range"; return false; }
    for(int64_t i = 0; i < 8 * H; ++i) if((int64_t)hexs[i] >= V) { err = "hex vertex id out of range"; return false; }
    for(int64_t i = 0; i < 6 * P; ++i) if((int64_t)pris[i] >= V) { err = "prism vertex id out of range"; return false; }
    if(4 * T + 6 * P + 8 * H >= (int64_t)1 << 32) { err = "too many element references"; return false; }
    return true;
}

// rec: H hexahedron records, then P prism records (makeRecords).
static void buildAdjacency(int64_t V, int64_t T, const uint32_t* tets, int64_t H, int64_t P, const uint32_t* rec,
                           Adjacency& a)
{
    // ---- Patch element lists (compileNeighborhoods: tets in id order, then
    // prisms, then hexes).
    std::vector<uint32_t>& eoff = a.eoff;
    eoff.assign(V + 1, 0);
    for(int64_t i = 0; i < 4 * T; ++i) ++eoff[tets[i] + 1];
    for(int64_t p = 0; p < P; ++p)
        for(int k = 0; k < 6; ++k) ++eoff[rec[8 * (H + p) + k] + 1];
    for(int64_t i = 0; i < 8 * H; ++i) ++eoff[rec[i] + 1];
    for(int64_t v = 0; v < V; ++v) eoff[v + 1] += eoff[v];
    std::vector<uint32_t>& eref = a.eref;
    eref.assign(eoff[V], 0);
    {
        std::vector<uint32_t> cur(eoff.begin(), eoff.end() - 1);
        for(int64_t e = 0; e < T; ++e)
            for(int k = 0; k < 4; ++k)
                eref[cur[tets[4 * e + k]]++] = ((uint32_t)e << 4) | (uint32_t)k;
        for(int64_t p = 0; p < P; ++p)
            for(int k = 0; k < 6; ++k)
                eref[cur[rec[8 * (H + p) + k]]++] = ((uint32_t)(H + p) << 4) | 8u | (uint32_t)k;
        for(int64_t e = 0; e < H; ++e)
            for(int k = 0; k < 8; ++k)
                eref[cur[rec[8 * e + k]]++] = ((uint32_t)e << 4) | 8u | (uint32_t)k;
    }

    // ---- Neighbour vertices (addEdge: append when absent, both directions).
    // Every element adds at most 3 edges at each of its vertices.
    std::vector<uint32_t>& ncnt = a.ncnt;
    std::vector<uint32_t>& ntmp = a.ntmp;
    ncnt.assign(V, 0);
    ntmp.assign((size_t)3 * eoff[V], 0);
    auto addEdge = [&](uint32_t a, uint32_t b) {
        uint32_t* la = &ntmp[(size_t)3 * eoff[a]];
        for(uint32_t i = 0; i < ncnt[a]; ++i)
            if(la[i] == b)
                return;
        la[ncnt[a]++] = b;
        ntmp[(size_t)3 * eoff[b] + ncnt[b]++] = a;
    };
    for(int64_t e = 0; e < T; ++e)
        for(int k = 0; k < 6; ++k)
            addEdge(tets[4 * e + TET_EDGES[k][0]], tets[4 * e + TET_EDGES[k][1]]);
    for(int64_t p = 0; p < P; ++p)
        for(int k = 0; k < 9; ++k)
            addEdge(rec[8 * (H + p) + PRI_EDGES[k][0]], rec[8 * (H + p) + PRI_EDGES[k][1]]);
    for(int64_t e = 0; e < H; ++e)
        for(int k = 0; k < 12; ++k)
            addEdge(rec[8 * e + HEX_EDGES[k][0]], rec[8 * e + HEX_EDGES[k][1]]);
}

// NodeGroups::determineTypes (NodeGroups.cpp:124-241): a vertex on a
// sliding constraint is a boundary vertex, a free vertex of an element that
// holds one is a subsurface vertex, any other free vertex is interior.
static bool computeTypes(int64_t V, const int32_t* fixed, const int32_t* cons, int64_t T, const uint32_t* tets,
                         int64_t H, int64_t P, const uint32_t* rec, const Adjacency& a, std::vector<uint8_t>& types,
                         std::string& err)
{
    types.assign(V, TYPE_INTERIOR);
    bool sliding = false;
    for(int64_t v = 0; v < V; ++v)
    {
        if(fixed[v])
        {
            if(cons && cons[v] >= 0) { err = "vertex " + std::to_string(v) + " is both fixed and sliding"; return false; }
            types[v] = TYPE_FIXED;
        }
        else if(cons && cons[v] >= 0)
        {
            types[v] = TYPE_BOUNDARY;
            sliding = true;
        }
    }
    if(!sliding) return true;
    std::vector<uint8_t> boundElem((size_t)(T + H + P), 0);
    for(int64_t e = 0; e < T; ++e)
        for(int k = 0; k < 4; ++k)
            if(types[tets[4 * e + k]] == TYPE_BOUNDARY) { boundElem[e] = 1; break; }
    for(int64_t e = 0; e < H + P; ++e)
        for(int k = 0; k < recordVerts(H, (uint32_t)e); ++k)
            if(types[rec[8 * e + k]] == TYPE_BOUNDARY) { boundElem[T + e] = 1; break; }
    for(int64_t v = 0; v < V; ++v)
    {
        if(types[v] != TYPE_INTERIOR) continue;
        for(uint32_t r = a.eoff[v]; r < a.eoff[v + 1]; ++r)
        {
            const uint32_t ref = a.eref[r], id = ref >> 4;
            if(boundElem[(ref & 8) ? T + id : id]) { types[v] = TYPE_SUBSURFACE; break; }
        }
    }
    return true;
}

// Orders each group's members by type (stable) and fills groupNodes,
// groupOff, runOff, runMaxPatch, groupMaxPatch and serialGroup from
// per-group member lists (ids of `out`'s numbering).
static void finishGroups(const std::vector<std::vector<uint32_t>>& members, Topology& out)
{
    const size_t G = members.size();
    out.groupOff.assign(G + 1, 0);
    out.runOff.assign(3 * G + 1, 0);
    out.runMaxPatch.assign(3 * G, 0);
    out.groupMaxPatch.assign(G, 0);
    out.groupNodes.clear();
    for(size_t g = 0; g < G; ++g)
    {
        for(int t = 0; t < 3; ++t)
        {
            out.runOff[3 * g + t] = (int64_t)out.groupNodes.size();
            for(uint32_t v : members[g])
                if(out.vtype[v] == TYPE_BOUNDARY + t)
                {
                    out.groupNodes.push_back(v);
                    const int32_t patch = (int32_t)(out.eoff[v + 1] - out.eoff[v]);
                    out.runMaxPatch[3 * g + t] = std::max(out.runMaxPatch[3 * g + t], patch);
                    out.groupMaxPatch[g] = std::max(out.groupMaxPatch[g], patch);
                }
        }
        out.groupOff[g + 1] = (int64_t)out.groupNodes.size();
    }
    out.runOff[3 * G] = (int64_t)out.groupNodes.size();
}

bool build(int64_t V, const double* pos, const int32_t* fixed, const int32_t* cons,
           int64_t T, const uint32_t* tets, int64_t H, const uint32_t* hexs,
           int64_t P, const uint32_t* pris, Topology& out, std::string& err)
{
    if(!checkInputs(V, T, tets, H, hexs, P, pris, err)) return false;
    const std::vector<uint32_t> rec = makeRecords(H, hexs, P, pris);
    Adjacency adj;
    buildAdjacency(V, T, tets, H, P, rec.data(), adj);
    const std::vector<uint32_t>& eoff = adj.eoff;
    const std::vector<uint32_t>& eref = adj.eref;
    const std::vector<uint32_t>& ncnt = adj.ncnt;
    const std::vector<uint32_t>& ntmp = adj.ntmp;

    // ---- Types (determineTypes).
    std::vector<uint8_t> types;
    if(!computeTypes(V, fixed, cons, T, tets, H, P, rec.data(), adj, types, err)) return false;

    // ---- Greedy colouring (determineGroups, NodeGroups.cpp:243-363): the
    // smallest existing colour unused by already-coloured element-mates,
    // else a new colour; breadth-first from the lowest unvisited id.
    std::vector<int> groups(V, NO_GROUP);
    std::vector<uint32_t> nextNodes;
    nextNodes.reserve(V);
    std::vector<uint64_t> usedBy;   // colour -> token of the last vertex that saw it
    int64_t seekStart = 0;
    while((int64_t)nextNodes.size() < V)
    {
        size_t firstNode = nextNodes.size();
        for(int64_t vId = seekStart; vId < V; ++vId)
        {
            ++seekStart;
            if(groups[vId] == NO_GROUP)
            {
                groups[vId] = UNSET_GROUP;
                nextNodes.push_back((uint32_t)vId);
                break;
            }
        }
        for(size_t i = firstNode; i < nextNodes.size(); ++i)
        {
            uint32_t vId = nextNodes[i];
            uint64_t token = (uint64_t)vId + 1;
            if(H == 0 && P == 0)
            {
                // Tetrahedra only: the element-mates are the edge neighbours,
                // and the neighbour list holds them in the order the element
                // walk below meets them (elements in patch order, a tet's
                // other vertices in local order), without repeats.
                const uint32_t* ln = &ntmp[(size_t)3 * eoff[vId]];
                for(uint32_t n = 0; n < ncnt[vId]; ++n)
                {
                    int& g = groups[ln[n]];
                    if(g == NO_GROUP) { g = UNSET_GROUP; nextNodes.push_back(ln[n]); }
                    else if(g != UNSET_GROUP) usedBy[g] = token;
                }
            }
            else for(uint32_t r = eoff[vId]; r < eoff[vId + 1]; ++r)
            {
                uint32_t ref = eref[r];
                uint32_t id = ref >> 4;
                const uint32_t* ev = (ref & 8) ? &rec[8 * (size_t)id] : &tets[4 * (size_t)id];
                int nv = (ref & 8) ? recordVerts(H, id) : 4;
                for(int n = 0; n < nv; ++n)
                {
                    int& g = groups[ev[n]];
                    if(g == NO_GROUP) { g = UNSET_GROUP; nextNodes.push_back(ev[n]); }
                    else if(g != UNSET_GROUP) usedBy[g] = token;
                }
            }
            int g = -1;
            for(size_t c = 0; c < usedBy.size(); ++c)
                if(usedBy[c] != token) { g = (int)c; break; }
            if(g < 0) { g = (int)usedBy.size(); usedBy.push_back(0); }
            groups[vId] = g;
        }
    }

    // ---- Ordering (determinePositions, NodeGroups.cpp:365-450): the same
    // comparator through std::sort, so ties resolve identically.
    // The comparator's outcomes as one key per vertex (all fixed vertices
    // tie; movable ones order by group, then patch size); std::sort's moves
    // depend only on the outcomes, so sorting (key, id) pairs gives the same
    // order as sorting ids with the comparator.
    std::vector<std::pair<uint64_t, int>> keyed(V);
    for(int64_t v = 0; v < V; ++v)
    {
        const uint64_t key = types[v] == TYPE_FIXED
                                 ? 0
                                 : ((uint64_t)types[v] << 60) | ((uint64_t)(uint32_t)groups[v] << 32) |
                                       (uint64_t)(eoff[v + 1] - eoff[v]);
        keyed[v] = std::make_pair(key, (int)v);
    }
    std::sort(keyed.begin(), keyed.end(),
              [](const std::pair<uint64_t, int>& a, const std::pair<uint64_t, int>& b) { return a.first < b.first; });
    std::vector<int> indices(V);
    for(int64_t p = 0; p < V; ++p) indices[p] = keyed[p].second;
    out.permutation.assign(V, 0);
    for(int64_t p = 0; p < V; ++p) out.permutation[indices[p]] = p;
    const std::vector<int64_t>& Perm = out.permutation;

    // ---- Renumbering (clusterNodes, NodeGroups.cpp:452-619).
    out.nverts = V; out.ntets = T; out.nhexs = H; out.npris = P;
    out.pos.resize(3 * V);
    out.fixed.resize(V);
    out.eoff.assign(V + 1, 0);
    out.noff.assign(V + 1, 0);
    for(int64_t p = 0; p < V; ++p)
    {
        int old = indices[p];
        out.pos[3 * p] = pos[3 * old]; out.pos[3 * p + 1] = pos[3 * old + 1]; out.pos[3 * p + 2] = pos[3 * old + 2];
        out.fixed[p] = fixed[old] ? 1 : 0;
        out.eoff[p + 1] = out.eoff[p] + (eoff[old + 1] - eoff[old]);
        out.noff[p + 1] = out.noff[p] + ncnt[old];
    }
    out.eref.resize(out.eoff[V]);
    out.nbr.resize(out.noff[V]);
    for(int64_t p = 0; p < V; ++p)
    {
        int old = indices[p];
        std::copy(eref.begin() + eoff[old], eref.begin() + eoff[old + 1], out.eref.begin() + out.eoff[p]);
        const uint32_t* ln = &ntmp[(size_t)3 * eoff[old]];
        for(uint32_t i = 0; i < ncnt[old]; ++i)
            out.nbr[out.noff[p] + i] = (uint32_t)Perm[ln[i]];
    }
    out.tets.resize(4 * T);
    for(int64_t i = 0; i < 4 * T; ++i) out.tets[i] = (uint32_t)Perm[tets[i]];
    out.hexs.resize(rec.size());
    for(size_t i = 0; i < rec.size(); ++i) out.hexs[i] = (uint32_t)Perm[rec[i]];

    // ---- Groups (clusterNodes): movable nodes are sorted by type, then
    // colour, so a group's nodes in ascending new id are its boundary, then
    // subsurface, then interior nodes (the order of its undispatchedNodes);
    // without sliding vertices each colour is one contiguous run of new ids.
    // Colours that only fixed vertices use below the largest movable colour
    // stay as empty groups.
    out.nfixed = 0;
    for(int64_t v = 0; v < V; ++v) if(fixed[v]) ++out.nfixed;
    out.vtype.resize(V);
    for(int64_t p = 0; p < V; ++p) out.vtype[p] = types[indices[p]];
    if(cons)
    {
        bool any = false;
        for(int64_t v = 0; v < V && !any; ++v) any = types[v] == TYPE_BOUNDARY;
        if(any)
        {
            out.cons.resize(V);
            for(int64_t p = 0; p < V; ++p) out.cons[p] = types[indices[p]] == TYPE_BOUNDARY ? cons[indices[p]] : -1;
        }
    }
    int maxGroup = -1;
    for(int64_t p = out.nfixed; p < V; ++p) maxGroup = std::max(maxGroup, groups[indices[p]]);
    const int G = maxGroup + 1;
    std::vector<std::vector<uint32_t>> members(G);
    out.serialGroup.resize(V - out.nfixed);
    for(int64_t p = out.nfixed; p < V; ++p)
    {
        members[groups[indices[p]]].push_back((uint32_t)p);
        out.serialGroup[p - out.nfixed] = (uint32_t)p;
    }
    finishGroups(members, out);
    return true;
}

bool buildGrouped(int64_t V, const double* pos, const int32_t* fixed, const int32_t* cons, int64_t T,
                  const uint32_t* tets, int64_t H, const uint32_t* hexs, int64_t P, const uint32_t* pris, int64_t G,
                  const int64_t* groupOff, const uint32_t* groupNodes, Topology& out, std::string& err)
{
    if(!checkInputs(V, T, tets, H, hexs, P, pris, err)) return false;
    if(G < 0) { err = "negative group count"; return false; }
    if(G > 0 && (!groupOff || (groupOff[G] > 0 && !groupNodes))) { err = "null group arrays"; return false; }
    if(G > 0 && groupOff[0] != 0) { err = "group_off[0] must be 0"; return false; }
    for(int64_t g = 0; g < G; ++g)
        if(groupOff[g + 1] < groupOff[g]) { err = "group offsets must not decrease"; return false; }
    const int64_t M = G > 0 ? groupOff[G] : 0;
    int64_t nfixed = 0;
    for(int64_t v = 0; v < V; ++v) if(fixed[v]) ++nfixed;
    if(M != V - nfixed)
    {
        err = "the groups must hold every movable vertex exactly once (" + std::to_string(M) + " group entries, " +
              std::to_string(V - nfixed) + " movable vertices)";
        return false;
    }
    const std::vector<uint32_t> rec = makeRecords(H, hexs, P, pris);
    Adjacency adj;
    buildAdjacency(V, T, tets, H, P, rec.data(), adj);
    std::vector<uint8_t> types;
    if(!computeTypes(V, fixed, cons, T, tets, H, P, rec.data(), adj, types, err)) return false;
    // Membership: each movable vertex in exactly one group, no fixed vertex.
    std::vector<int32_t> groupOf(V, -1);
    for(int64_t g = 0; g < G; ++g)
        for(int64_t i = groupOff[g]; i < groupOff[g + 1]; ++i)
        {
            const uint32_t v = groupNodes[i];
            if((int64_t)v >= V) { err = "group vertex id out of range"; return false; }
            if(fixed[v]) { err = "fixed vertex " + std::to_string(v) + " is in group " + std::to_string(g); return false; }
            if(groupOf[v] >= 0) { err = "vertex " + std::to_string(v) + " is in two groups"; return false; }
            groupOf[v] = (int32_t)g;
        }
    // Independence (NodeGroups' invariant): no two vertices of one group share
    // an element.
    for(int64_t e = 0; e < T + H + P; ++e)
    {
        const bool record = e >= T;
        const uint32_t* ev = record ? &rec[8 * (size_t)(e - T)] : &tets[4 * (size_t)e];
        const int nv = record ? recordVerts(H, (uint32_t)(e - T)) : 4;
        for(int a = 0; a < nv; ++a)
            for(int b = a + 1; b < nv; ++b)
                if(ev[a] != ev[b] && groupOf[ev[a]] >= 0 && groupOf[ev[a]] == groupOf[ev[b]])
                {
                    const char* what = !record ? "tetrahedron " : e - T < H ? "hexahedron " : "prism ";
                    const int64_t id = !record ? e : e - T < H ? e - T : e - T - H;
                    err = "vertices " + std::to_string(ev[a]) + " and " + std::to_string(ev[b]) + " of group " +
                          std::to_string(groupOf[ev[a]]) + " share " + what + std::to_string(id);
                    return false;
                }
    }
    out.nverts = V; out.ntets = T; out.nhexs = H; out.npris = P;
    out.pos.assign(pos, pos + 3 * V);
    out.fixed.resize(V);
    for(int64_t v = 0; v < V; ++v) out.fixed[v] = fixed[v] ? 1 : 0;
    out.tets.assign(tets, tets + 4 * T);
    out.hexs = rec;
    out.eoff = adj.eoff;
    out.eref = adj.eref;
    out.noff.assign(V + 1, 0);
    for(int64_t v = 0; v < V; ++v) out.noff[v + 1] = out.noff[v] + adj.ncnt[v];
    out.nbr.resize(out.noff[V]);
    for(int64_t v = 0; v < V; ++v)
        std::copy(adj.ntmp.begin() + (size_t)3 * adj.eoff[v], adj.ntmp.begin() + (size_t)3 * adj.eoff[v] + adj.ncnt[v],
                  out.nbr.begin() + out.noff[v]);
    out.permutation.resize(V);
    std::iota(out.permutation.begin(), out.permutation.end(), (int64_t)0);
    out.nfixed = nfixed;
    out.vtype = types;
    if(cons)
        for(int64_t v = 0; v < V; ++v)
            if(types[v] == TYPE_BOUNDARY)
            {
                out.cons.resize(V);
                for(int64_t u = 0; u < V; ++u) out.cons[u] = types[u] == TYPE_BOUNDARY ? cons[u] : -1;
                break;
            }
    std::vector<std::vector<uint32_t>> members((size_t)G);
    for(int64_t g = 0; g < G; ++g) members[g].assign(groupNodes + groupOff[g], groupNodes + groupOff[g + 1]);
    finishGroups(members, out);
    // Serial order (NodeGroups::serialGroup): type-major, then colour.
    out.serialGroup.clear();
    for(int t = 0; t < 3; ++t)
        for(int64_t g = 0; g < G; ++g)
            out.serialGroup.insert(out.serialGroup.end(), out.groupNodes.begin() + out.runOff[3 * g + t],
                                   out.groupNodes.begin() + out.runOff[3 * g + t + 1]);
    return true;
}

} // namespace gmt
