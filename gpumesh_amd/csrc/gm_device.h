// gm_device.h -- the interface between the host runtime (api.hip) and the
// precision-specific kernel translation units (kernels_f64.hip, kernels_f32.hip).
#ifndef GM_DEVICE_H
#define GM_DEVICE_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "gm_quality.h"

namespace gmd
{

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
    const uint4* hexs;        // 2 per hex (v0..v3, v4..v7)
    const uint32_t* eoff;     // V+1
    const uint32_t* eref;
    const uint32_t* noff;     // V+1
    const uint32_t* nbr;
    const uint8_t* fixed;     // 1 = fixed vertex (never moved; snaps to itself)
    uint32_t nverts, ntets, nhexs;
};

enum { SM_NELDER_MEAD = 0, SM_GRADIENT_DESCENT = 1, SM_SPAWN_SEARCH = 2, SM_GETME = 3 };
enum { KS_PLAIN = 0, KS_WAVE = 1 };

// Scratch for the mesh-quality reduction.
struct EvalScratch
{
    double* partials;   // 3 per block: min, sum, compensation
    int32_t* flags;     // per block: 1 when an element quality <= 0 (sum is +inf)
    double* result;     // [min, invSum, n] per recorded pass (stride 3)
    int32_t* hist;      // 20 running bucket counts (reset by the final kernel)
    int32_t* histOut;   // 20 per recorded pass
    double* elemQual;   // optional per-element qualities (nullptr = off)
};

const int EVAL_BLOCKS = 1024;
const int EVAL_THREADS = 256;

// Kernel launchers of one (precision T, evaluator EV, sampler SA)
// combination.  All launches go on `stream`; nothing here synchronises.
template<typename T, int EV, int SA>
struct Launchers
{
    static hipError_t smooth(const DevMesh& m, const gmq::Crew& c, int smoother, int shape,
                             const uint32_t* nodes, uint32_t count, double moveCoeff,
                             int32_t* evalCounts, int maxPatch, hipStream_t stream);
    static hipError_t getmeElements(const DevMesh& m, const gmq::Crew& c, void* tetOut, void* hexOut,
                                    hipStream_t stream);
    static hipError_t getmeVertices(const DevMesh& m, const gmq::Crew& c, const uint32_t* nodes, uint32_t count,
                                    const void* tetOut, const void* hexOut, int32_t* evalCounts,
                                    hipStream_t stream);
};

// FP64 mesh-quality reduction of one (evaluator, sampler) combination, reading
// the FP64 (evaluateF64) or FP32 (evaluateF32) position copy.  Always built
// without contraction, with the FP64 kernels.
template<int EV, int SA>
struct Evaluators
{
    static hipError_t evaluateF64(const DevMesh& m, const gmq::Crew& c, const EvalScratch& s, int pass,
                                  hipStream_t stream);
    static hipError_t evaluateF32(const DevMesh& m, const gmq::Crew& c, const EvalScratch& s, int pass,
                                  hipStream_t stream);
};

// Run-time selection over (evaluator, sampler) (dispatch.hip).
template<typename T>
hipError_t launchSmoothGroup(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                             int smoother, int shape, const uint32_t* nodes, uint32_t count,
                             double moveCoeff, int32_t* evalCounts, int maxPatch, hipStream_t stream);

template<typename T>
hipError_t launchGetmeElements(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                               void* tetOut, void* hexOut, hipStream_t stream);

template<typename T>
hipError_t launchGetmeVertices(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                               const uint32_t* nodes, uint32_t count, const void* tetOut,
                               const void* hexOut, int32_t* evalCounts, hipStream_t stream);

// Mesh quality in FP64 from the positions stored in precision P (double:
// pos64, float: pos32); writes slot `pass` of scratch.result / scratch.histOut.
template<typename P>
hipError_t launchEvaluate(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                          const EvalScratch& s, int pass, hipStream_t stream);

// Position precision conversions.
hipError_t launchToF32(const DevMesh& m, hipStream_t stream);
hipError_t launchToF64(const DevMesh& m, hipStream_t stream);

// Math self-test: op 0 sqrt, 1 div (a/b), 2 pow23, 3 cbrt, 4 exp; n values.
hipError_t launchMathProbe(int op, const double* a, const double* b, double* out, int n, hipStream_t stream);

} // namespace gmd

#endif // GM_DEVICE_H
