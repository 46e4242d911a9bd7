// gm_batr.h -- the BATR topologist on the host: face swaps, edge swaps and
// edge splits/merges of a tetrahedral mesh (Topologists/BatrTopologist.cpp),
// between the device smoothing passes of a global pass.  Not part of the C
// ABI (gm_mesh_restructure in api.hip is its entry point).
#ifndef GM_BATR_H
#define GM_BATR_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "gm_quality.h"

namespace gmb
{

// A tetrahedral mesh in the numbering of a compiled topology, with that
// topology's patch lists (eoff/eref, eref >> 4 = tet id) and neighbour lists
// (noff/nbr) in the reference's order: the operations read them in that
// order, so their results depend on it.
struct TetMesh
{
    std::vector<double> pos;       // 3 per vertex
    std::vector<uint8_t> fixed;    // 1: fixed vertex
    std::vector<uint32_t> tets;    // 4 per tet
    std::vector<uint32_t> eoff, eref, noff, nbr;
};

struct Settings
{
    int topoPasses = 5;          // Schedule::topoOperationPassCount
    int refinementSweeps = 5;    // Schedule::refinementSweepCount
    double minEdge = 0.5;        // AbstractTopologist::minEdgeLength (in the crew's metric)
    double maxEdge = 1.5;        // AbstractTopologist::maxEdgeLength
    int evaluator = 0;           // GM_MEAN_RATIO / GM_METRIC_CONFORMITY
    int sampler = 0;             // GM_UNIFORM ... GM_ANALYTIC (crew.texels on the host)
    gmq::Crew crew{};
};

struct Counts
{
    int64_t faceSwaps = 0, edgeSwaps = 0, splits = 0, merges = 0, passes = 0;
};

// Runs the operations on `in` and writes the resulting vertices (positions,
// fixed flags) and tets to `out` (its list fields are left empty: the caller
// recompiles the topology).  Vertices must be free or fixed.  Returns false
// with `err` set when the input is not a valid tetrahedral mesh.
bool restructure(const TetMesh& in, const Settings& s, TetMesh& out, Counts& counts, std::string& err);

} // namespace gmb

#endif // GM_BATR_H
