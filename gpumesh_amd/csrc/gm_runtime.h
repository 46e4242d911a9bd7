// gm_runtime.h -- host runtime internals shared by runtime.hip (plans, pass
// loop) and api.hip (the C ABI).  Not part of the C ABI.
#ifndef GM_RUNTIME_H
#define GM_RUNTIME_H

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/gpumesh.h"
#include "gm_device.h"
#include "gm_topology.h"

namespace gmr
{

// Records the message for gm_last_error() and returns `code`.
int fail(int code, const std::string& msg);
// Message of the calling thread's last failure.
const char* lastError();

// A vertex work list of one colour step, element kind and patch class.
struct VList
{
    const uint4* segs = nullptr;   // device spans (one per member with vertices in the list)
    uint32_t nsegs = 0;
    uint32_t count = 0;
    int kind = gmd::EK_TET;
    int maxPatch = 0;
    int constrained = 0;           // vertices sliding on constraints
    size_t firstSpan = 0;          // offset of segs in the plan's span buffer
};

struct RunParams
{
    int smoother = 0;
    int passes = 0;
    int precision = GM_FP64;
    int polishPasses = 1;        // FP64 passes ending a GM_FP32_POLISH run
    int kernel = GM_KERNEL_WAVE;
    double spawnCoeff = 1.0 / 20.0;
    gmd::LaneShape lane{1, 0};   // tileElems 0: chosen per work list
    bool graph = false;
    bool timeKernels = true;   // per-launch events (not inside graphs)
    // Autopilot (at most `passes` sweeps; see DevMesh::autopilot).
    bool autopilot = false;
    double apMeanGain = 0.001;   // Schedule::minQualThreshold
    double apMinGain = 0.0;      // Schedule::qualMeanThreshold
};

struct Plan;

// Pass `pass` of a run of `passes` sweeps runs in FP64 ...
inline bool passF64(int precision, int pass, int passes, int polishPasses)
{
    return precision == GM_FP64 || precision == GM_FP64_FAST ||
           (precision == GM_FP32_POLISH && pass >= passes - polishPasses);
}
// ... and, when it does, on the mixed-precision FP64 kernels unless the run
// is exact FP64.
inline bool passFast(int precision) { return precision == GM_FP32_POLISH || precision == GM_FP64_FAST; }
// A run starts from the FP32 copy of the positions.
inline bool startsF32(int precision) { return precision == GM_FP32 || precision == GM_FP32_POLISH; }

// Process-wide number of FP64 passes ending a GM_FP32_POLISH run.
int polishPasses();
void setPolishPasses(int n);

} // namespace gmr

// The read-only device arrays of one topology on one device (elements,
// patch and neighbour lists, fixed flags, sliding constraint ids, slot map,
// GETMe corner slots), shared by every mesh of that topology there
// (runtime.hip, deviceTopoOf).  The destructor returns them to the pool.
struct DeviceTopo
{
    uint4* tets = nullptr;
    uint4* hexs = nullptr;
    uint32_t *eoff = nullptr, *eref = nullptr, *noff = nullptr, *nbr = nullptr, *slotOf = nullptr;
    uint8_t* fixed = nullptr;
    int32_t* cons = nullptr;
    uint32_t* cornerSlot = nullptr;   // made by the first GETMe run on this device
    DeviceTopo() = default;
    DeviceTopo(const DeviceTopo&) = delete;
    DeviceTopo& operator=(const DeviceTopo&) = delete;
    ~DeviceTopo();
};

// Device slots of a topology's vertices and the per-slot arrays uploaded for
// it (runtime.hip, buildLayout).
struct DeviceLayout
{
    // Device copies, one per device, made by the first upload there and kept
    // as long as the layout lives: with its topology while the topology cache
    // holds it, else with the meshes that use it.  devMu guards them.
    mutable std::mutex devMu;
    mutable std::map<int, std::shared_ptr<DeviceTopo>> onDevice;

    std::vector<uint32_t> slotOf;     // reference number -> slot
    std::vector<uint32_t> vertexAt;   // slot -> reference number
    std::vector<uint32_t> eoff, eref, noff, nbr, tets, hexs;
    std::vector<uint8_t> fixed;
    std::vector<int32_t> cons;        // sliding constraint per slot (empty when none slides)
    std::vector<uint32_t> cornerSlot; // patch entry of each element corner
};

struct gm_mesh
{
    bool onDevice = false;     // false: topology only (gm_mesh_create_host)
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<const gmt::Topology> topo;   // shared by meshes of equal connectivity
    std::vector<double> hostPos;  // renumbered input positions, until uploaded
    // Device slot of each vertex (reference numbering -> slot): its input id.
    // Per-vertex device arrays are laid out by slot, so vertices that are
    // close in the input numbering (mesh generators and files number them
    // with spatial coherence) are close in memory; the reference numbering
    // scatters each colour group over the whole mesh.
    std::vector<uint32_t> slotOf;
    std::shared_ptr<const DeviceLayout> layout;   // the topology's device layout (slotOf is its copy)
    std::shared_ptr<DeviceTopo> devTopo;          // its device arrays, shared (dm points into them)
    uint32_t* dSlotOf = nullptr;     // device copy of slotOf
    int32_t* dCountsRef = nullptr;   // evaluation counts in the reference numbering
    gmd::DevMesh dm{};         // host copy of the device descriptor (owns its pointers)
    int kind = gmd::EK_TET;    // element kind of the whole mesh
    int maxPatch = 0;
    int resultCap = 0;
    bool getmeReady = false;
    double* elemQualBuf = nullptr;
    int evaluator = GM_MEAN_RATIO;
    int sampler = GM_UNIFORM;
    gmq::Crew crew{};
    double aspect = 1.0;      // aspect ratio of the GM_ANALYTIC field
    // Sliding constraints: gmd::CONS_STRIDE doubles per entry (host copy;
    // the device copy is dm.consTable).
    std::vector<double> consTable;
    float* texels = nullptr;
    uint64_t texHash = 0;      // FNV-1a of the texels (batch members must share the grid)
    gmr::Plan* plan = nullptr; // the mesh as a batch of one
};

namespace gmr
{

struct Plan
{
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<gm_mesh*> meshes;
    gmd::DevMesh* dTable = nullptr;
    std::vector<std::vector<VList>> colours;      // [colour][list]: vertex-wise sweeps, colour by colour
    // GETMe's vertex steps in NodeGroups' serial order (boundary runs of every
    // colour, then subsurface, then interior); empty when no member has
    // sliding vertices (then they follow `colours`).
    std::vector<std::vector<VList>> getmeSteps;
    gmd::BlockWork tetBlocks{}, hexBlocks{}, priBlocks{}, evalBlocks{}, vertBlocks{};
    uint32_t* meshList = nullptr; uint32_t nMeshes = 0;
    uint4* spans = nullptr;                       // every span table of the plan, one buffer
    std::vector<void*> owned;                     // device allocations of this plan
    // Graphs of one pass body, keyed by (smoother, fp64, kernel, ppg, tile, spawn coeff bits, eval f32).
    std::map<std::tuple<int, int, int, int, int, uint64_t, int>, hipGraphExec_t> graphs;
    std::vector<hipGraph_t> graphDefs;
    // Timing.
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> pool;
    std::vector<int> poolTag;
    size_t poolUsed = 0;
    double totalMs = 0.0;
    double kernMs[GM_KT_COUNT] = {0};
    int64_t kernLaunches[GM_KT_COUNT] = {0};
    // Graphs hold the crew by value: they are dropped when it changes.
    gmq::Crew graphCrew{};
    int graphEv = -1, graphSa = -1;
    // Host copy of the descriptor table last uploaded.
    std::vector<gmd::DevMesh> table;
};

void clearGraphs(Plan* p);

// Caching device allocator (runtime.hip): poolFree of a pointer that did not
// come from poolAlloc falls back to hipFree.
hipError_t poolAlloc(void** out, size_t bytes);
void poolFree(void* p);
void setPoolCap(size_t bytes);
size_t poolCachedBytes();

// Host topology of a new handle (through the topology cache).
// npris prisms (6 vertices each) may follow.
int makeMesh(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
             int64_t nhexs, const uint32_t* hexs, gm_mesh** out, const int32_t* cons = nullptr,
             const std::vector<double>* consTable = nullptr, int64_t npris = 0, const uint32_t* pris = nullptr);
// The GM_ANALYTIC field's parameters of a crew: the aspect ratio and the
// sampler's fixed rotation (AbstractSampler's constructor,
// Samplers/AbstractSampler.cpp:30-41) and its transpose.
void setAnalyticField(gmq::Crew& c, double aspect);
// FNV-1a (64-bit) over the bytes of n texels: the identity of a texture grid
// for crews that must share one.
uint64_t texelHash(const float* texels, size_t n);

// makeMesh with the caller's numbering and colour groups (gmt::buildGrouped).
int makeMeshGrouped(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
                    int64_t nhexs, const uint32_t* hexs, int64_t ngroups, const int64_t* groupOff,
                    const uint32_t* groupNodes, gm_mesh** out, const int32_t* cons = nullptr,
                    const std::vector<double>* consTable = nullptr, int64_t npris = 0,
                    const uint32_t* pris = nullptr);
// Device memory a mesh needs when resident (connectivity, positions, reduction
// buffers, `records` pass records, GETMe scratch when `getme`).
size_t residentBytes(const gm_mesh* m, int records, bool getme);
// Allocates and uploads the device copy on the current device, with its own
// stream and a plan of one.  On failure everything allocated is released.
int uploadMesh(gm_mesh* m);
// Makes (or finds) the device copy of m's topology arrays on the current
// device without uploading the mesh itself.
int prepareDeviceTopology(const gm_mesh* m);
void destroyMesh(gm_mesh* m);

// Builds the fused plan of `meshes` (all resident on one device) on `stream`.
int buildPlan(const std::vector<gm_mesh*>& meshes, hipStream_t stream, Plan** out);
// Whether some member has hexahedra or prisms: the evaluation then runs the
// kernel with the record branch (the tet-only one reads every element as a
// tetrahedron).
bool planHasRecords(const Plan* p);
void freePlan(Plan* p);
// Makes sure every mesh has room for `records` pass records and (for GETMe)
// its element scratch, then uploads the descriptors.
int preparePlan(Plan* p, int records, bool getme);
// Runs the pass loop of `r` over the plan; synchronises at the end.
int runPlan(Plan* p, const gmq::Crew& crew, int evaluator, int sampler, const RunParams& r);
// One sweep restricted to colours [g0, g1) (no evaluation), FP64 or FP32.
int sweepColours(Plan* p, const gmq::Crew& crew, int evaluator, int sampler, const RunParams& r, int64_t g0,
                 int64_t g1);
// Pieces of the pass loop for the serving loop, launched on the plan's stream
// without synchronising: the run start (counters reset, FP32 copy when
// `toF32`, record 0), one pass (FP64 widening first when `convertFirst`; the
// reduction reads the FP32 copy unless `f64`), and the final widening.
// The plan must have been prepared (preparePlan) for its members.
int launchBegin(Plan* p, const gmq::Crew& crew, int evaluator, int sampler, bool toF32);
int launchPass(Plan* p, const gmq::Crew& crew, int evaluator, int sampler, const RunParams& r, bool f64,
               bool convertFirst);
int launchToF64All(Plan* p);
// Copies the pass records of mesh i (stats 3 per record, hist 20 per record).
int fetchRecords(Plan* p, size_t i, int records, double* stats, int32_t* hist, int32_t* evalCounts);
// Pass records mesh i holds after the last run (sweeps run + 1).
int recordsWritten(Plan* p, size_t i, int* records);

} // namespace gmr

#endif // GM_RUNTIME_H
