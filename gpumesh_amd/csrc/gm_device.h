// gm_device.h -- the interface between the host runtime (runtime.hip, api.hip)
// and the kernel translation units (kernels_inst.hip, dispatch.hip).
//
// Every launch is a fused launch over a table of resident meshes: a single
// mesh is a batch of one.  Work is described by span tables, one span per
// member: vertex work of a colour step is the concatenation of the members'
// colour groups, block work (element, evaluation and vertex blocks) the
// concatenation of the members' blocks.  A thread finds its span by binary
// search, so the tables stay O(members) however large the meshes are, and a
// new batch is planned without touching per-vertex data.  Meshes never share
// vertices, so any union of same-colour groups is an independent set.
#ifndef GM_DEVICE_H
#define GM_DEVICE_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "gm_quality.h"

namespace gmd
{

const int EVAL_MAX_BLOCKS = 2048;
const int EVAL_THREADS = 256;
const int ELEM_BLOCK = 256;        // elements per block of the GETMe element kernels

// Number of evaluation blocks of a mesh with E elements.  The same value is
// used alone and in a batch, so a mesh's min/mean reduction tree (and hence
// its result bits) does not depend on what it is batched with.
inline uint32_t evalBlocksFor(uint64_t E)
{
    uint64_t b = (E + 1023) / 1024;
    if(b < 1) b = 1;
    if(b > (uint64_t)EVAL_MAX_BLOCKS) b = EVAL_MAX_BLOCKS;
    return (uint32_t)b;
}

// Resident mesh, as the kernels see it.  Positions are 16/32-byte AoS
// (x, y, z, pad) so that a vertex gather is one aligned vector load.
// Element references of a vertex's patch are packed as
// (element id << 4) | (is_hex << 3) | local vertex index, in the order of
// Mesh::compileNeighborhoods (tets by ascending id, then hexes).
struct DevMesh
{
    double4* pos64;
    float4* pos32;
    const uint4* tets;        // 1 per tet
    const uint4* hexs;        // 2 per record: nhexs hexahedra (v0..v3, v4..v7), then npris prisms
                              // (v0..v3, v4 v5 v5 v5)
    const uint32_t* eoff;     // V+1
    const uint32_t* eref;
    const uint32_t* noff;     // V+1
    const uint32_t* nbr;
    const uint8_t* fixed;     // 1 = fixed vertex (never moved; snaps to itself)
    int32_t* evalCounts;      // patch evaluations per vertex (accumulated)
    void* getmePull;          // GETMe (vpp, weight) per patch entry (eref order), sized for FP64
    const uint32_t* cornerSlot; // patch entry of each element corner: tets 4 e + k, then hexes 4 T + 8 h + k
    void* getmeQual;          // GETMe: current quality per element (tets, then hexes), sized for FP64
    // Quality reduction of this mesh.
    double* evalPartials;     // 3 per block: min, sum, compensation
    int32_t* evalFlags;       // per block: 1 when an element quality <= 0
    int32_t* evalHist;        // 20 running bucket counts (reset by the final kernel)
    double* evalResult;       // [min, invSum, n] per recorded pass
    int32_t* evalHistOut;     // 20 per recorded pass
    double* elemQual;         // optional per-element qualities (nullptr = off)
    int32_t* passSlot;        // [0] next pass record to write (advanced by the final kernel),
                              // [1] 1 once the autopilot stopped the run (pass kernels then do nothing)
    uint32_t nverts, ntets, nhexs, npris;
    uint32_t evalBlocks;
    uint32_t resultCap;       // pass records evalResult / evalHistOut hold
    // Pass schedule (AbstractSmoother::evaluateMeshQuality,
    // Smoothers/AbstractSmoother.cpp:186-197): with autopilot set, the record
    // k >= 1 stops the run unless meanGain > apMeanGain or minGain > apMinGain
    // (Schedule::minQualThreshold and Schedule::qualMeanThreshold, which the
    // reference compares with those gains in that order).
    int32_t autopilot;
    double apMeanGain;
    double apMinGain;
    // Sliding constraints (nullptr when no vertex slides): per slot an index
    // into consTable or -1; CONS_STRIDE doubles per constraint.
    const int32_t* cons;
    const double* consTable;
};

// A constraint table entry: [kind, a.xyz, n.xyz, b.xyz, length, 0] -- a plane
// through a with unit normal n, or a segment from a to b with unit direction
// n and length `length`.
constexpr int CONS_STRIDE = 12;
enum { CONS_PLANE = 1, CONS_SEGMENT = 2 };

// True once the autopilot stopped this mesh's run: the pass kernels leave.
__device__ __forceinline__ bool runStopped(const DevMesh& m)
{
    return m.autopilot && m.passSlot[1] != 0;
}

enum { SM_NELDER_MEAD = 0, SM_GRADIENT_DESCENT = 1, SM_SPAWN_SEARCH = 2, SM_GETME = 3 };
enum { KS_PLAIN = 0, KS_WAVE = 1 };
// Element kind of a work list: the candidate-parallel kernels stage patches
// of one kind (4 or 8 slots per element); mixed lists use the wave kernel.
enum { EK_TET = 0, EK_HEX = 1, EK_MIXED = 2 };

// Launch shape of the candidate-parallel kernels: patches (vertices) per
// workgroup and the LDS tile (elements staged per workgroup).
struct LaneShape
{
    int patchesPerGroup;   // wavefronts per block: 1, 2, 4, 8 (spawn: 1 vertex each, GD: 8 vertices each,
                           // GETMe vertex update: 64 vertices each)
    int tileElems;         // LDS elements staged per wavefront: 64, 128, 256 (spawn, GD)
};

// Span (x, y, z, w) = (mesh index, first group node, first index of the
// span in the launch, items in the span).  Block spans use y = 0 and
// z = first block.  Spans are sorted by z; index i belongs to the last span
// with z <= i.  Vertex spans start at multiples of ITEM_ALIGN, so the items
// of one wavefront of a wave or lane kernel (at most 16) lie in one span; the
// indices between a span's last item and the next span are holes.
constexpr uint32_t ITEM_ALIGN = 16;
constexpr uint32_t NO_ITEM = 0xffffffffu;
__host__ __device__ inline uint32_t spanOf(const uint4* spans, uint32_t nspans, uint32_t i)
{
    uint32_t lo = 0, hi = nspans - 1;
    while(lo < hi)
    {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if(spans[mid].z <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// One fused vertex launch.
struct VertexWork
{
    const DevMesh* meshes;     // device table
    const uint4* segs;         // spans over the members' colour groups
    uint32_t nsegs;
    uint32_t count;            // vertices in the launch
    int kind;                  // EK_*
    int maxPatch;              // largest patch in the list
    int polish;                // FP64 polish pass of the mixed-precision mode (not the exact FP64 mode)
    int constrained;           // the list's vertices slide on constraints (generic kernels only)
};

// Block work: `count` blocks spread over `nspans` members.
struct BlockWork
{
    const uint4* spans;
    uint32_t nspans;
    uint32_t count;
};

// Kernel launchers of one (precision T, evaluator EV, sampler SA)
// combination.  All launches go on `stream`; nothing here synchronises.
template<typename T, int EV, int SA>
struct Launchers
{
    static hipError_t smooth(const VertexWork& w, const gmq::Crew& c, int smoother, int shape,
                             LaneShape lane, double moveCoeff, hipStream_t stream);
    // fast: the mixed-precision FP64 passes (seeded routines for mean ratio
    // with the uniform sampler on tets).
    static hipError_t getmeElements(const DevMesh* meshes, BlockWork tets, BlockWork hexs, BlockWork pris,
                                    const gmq::Crew& c, bool fast, hipStream_t stream);
    // GETMe vertex updates: KS_WAVE, one vertex per group of lanes (patches
    // of up to 64 elements); otherwise one vertex per thread, 64 x
    // lane.patchesPerGroup threads per block.
    static hipError_t getmeVertices(const VertexWork& w, const gmq::Crew& c, int shape, LaneShape lane,
                                    hipStream_t stream);
};

// FP64 mesh-quality reduction of one (evaluator, sampler) combination, reading
// the FP64 (evaluateF64) or FP32 (evaluateF32) position copy.  `blocks` spans
// the members' evaluation blocks; `meshList` the meshes to finish.  Always
// built without contraction, with the FP64 kernels.
template<int EV, int SA>
struct Evaluators
{
    static hipError_t evaluateF64(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                  uint32_t nmeshes, const gmq::Crew& c, bool hexes, hipStream_t stream);
    static hipError_t evaluateF32(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                  uint32_t nmeshes, const gmq::Crew& c, bool hexes, hipStream_t stream);
};

// Run-time selection over (evaluator, sampler) (dispatch.hip).
template<typename T>
hipError_t launchSmooth(const VertexWork& w, const gmq::Crew& c, int evaluator, int sampler, int smoother,
                        int shape, LaneShape lane, double moveCoeff, hipStream_t stream);
template<typename T>
hipError_t launchGetmeElements(const DevMesh* meshes, BlockWork tets, BlockWork hexs, BlockWork pris,
                               const gmq::Crew& c, int evaluator, int sampler, bool fast, hipStream_t stream);
template<typename T>
hipError_t launchGetmeVertices(const VertexWork& w, const gmq::Crew& c, int evaluator, int sampler, int shape,
                               LaneShape lane, hipStream_t stream);
// P = position precision read (double: pos64, float: pos32); always FP64 math.
// `hexes`: some mesh of the batch has hexahedra or prisms (a tet-only batch
// runs a kernel without the record branch, which needs fewer registers).
template<typename P>
hipError_t launchEvaluate(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList, uint32_t nmeshes,
                          const gmq::Crew& c, int evaluator, int sampler, bool hexes, hipStream_t stream);

// Position precision conversions over blocks of VERT_BLOCK vertices.
const int VERT_BLOCK = 256;
hipError_t launchToF32(const DevMesh* meshes, BlockWork vertexBlocks, hipStream_t stream);
// out[v] = counts[slotOf[v]], v < n (per-vertex values from the device layout
// to the reference numbering).
hipError_t launchGatherCounts(const int32_t* counts, const uint32_t* slotOf, int32_t* out, uint32_t n,
                              hipStream_t stream);
hipError_t launchToF64(const DevMesh* meshes, BlockWork vertexBlocks, hipStream_t stream);
// Zeroes evalCounts and sets passSlot = 0 for the meshes of a batch (fill
// kernels, so that the reset can be captured in a graph).
hipError_t launchResetRun(const DevMesh* meshes, BlockWork vertexBlocks, const uint32_t* meshList,
                          uint32_t nmeshes, hipStream_t stream);

// True when the candidate-parallel kernel of `smoother` runs `w` with `lane`
// (otherwise the wave kernel does).
bool laneFits(const VertexWork& w, int smoother, LaneShape lane, bool fp64, int evaluator, int sampler);

// Math self-test: op 0 sqrt, 1 div (a/b), 2 pow23, 3 cbrt, 4 exp; n values.
hipError_t launchMathProbe(int op, const double* a, const double* b, double* out, int n, hipStream_t stream);
hipError_t launchThroughput(int op, int blocks, int iters, float* out, hipStream_t stream);
int throughputChains();

} // namespace gmd

#endif // GM_DEVICE_H
