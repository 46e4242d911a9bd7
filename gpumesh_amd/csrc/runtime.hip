// runtime.hip -- fused execution plans over resident meshes and the pass
// loop of AbstractSmoother::smoothMesh (Smoothers/AbstractSmoother.cpp:57-211)
// on the device.  A mesh handle runs as a plan of one; a batch is a plan of
// many.  Everything a plan launches reads the per-mesh descriptors from a
// device table, so a mesh's kernels, launch order within the mesh and
// reduction trees are the same alone and batched.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>

#include <algorithm>
#include <cstring>
#include <map>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "gm_runtime.h"

namespace gmr
{

static thread_local std::string g_lastError;

int fail(int code, const std::string& msg)
{
    g_lastError = msg;
    return code;
}

const char* lastError() { return g_lastError.c_str(); }

#define GMR_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if(e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP,         \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while(0)

// ---------------------------------------------------------------------------
// Caching device allocator.  Freed blocks are kept per (device, size class)
// and handed out again, so that meshes, pass records and plans that come and
// go in a serving loop cost neither hipMalloc latency nor the device-wide
// synchronisation of hipFree.  Size classes are 256 B, then four per power of
// two (at most 25 % rounding).  Blocks are only freed or reused after the
// streams that used them have been synchronised by their owners.
// ---------------------------------------------------------------------------
namespace
{
std::mutex g_poolMutex;
std::map<std::pair<int, size_t>, std::vector<void*>> g_poolFree;
std::map<void*, std::pair<int, size_t>> g_poolLive;
size_t g_poolCached = 0, g_poolCap = (size_t)16 << 30;

size_t sizeClass(size_t bytes)
{
    if(bytes <= 256) return 256;
    int e = 63 - __builtin_clzll((unsigned long long)bytes);
    const size_t step = (size_t)1 << (e - 2);
    return (bytes + step - 1) / step * step;
}
} // namespace

hipError_t poolAlloc(void** out, size_t bytes)
{
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if(e != hipSuccess) return e;
    const size_t sz = sizeClass(bytes ? bytes : 1);
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        auto it = g_poolFree.find({dev, sz});
        if(it != g_poolFree.end() && !it->second.empty())
        {
            *out = it->second.back();
            it->second.pop_back();
            g_poolCached -= sz;
            g_poolLive[*out] = {dev, sz};
            return hipSuccess;
        }
    }
    e = hipMalloc(out, sz);
    if(e == hipErrorOutOfMemory)
    {
        // Give the cached blocks of this device back and try once more.
        (void)hipGetLastError();
        std::vector<void*> drop;
        {
            std::lock_guard<std::mutex> lock(g_poolMutex);
            for(auto& kv : g_poolFree)
                if(kv.first.first == dev)
                {
                    for(void* q : kv.second) { drop.push_back(q); g_poolCached -= kv.first.second; }
                    kv.second.clear();
                }
        }
        for(void* q : drop) (void)hipFree(q);
        e = hipMalloc(out, sz);
    }
    if(e != hipSuccess)
    {
        *out = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> lock(g_poolMutex);
    g_poolLive[*out] = {dev, sz};
    return hipSuccess;
}

void poolFree(void* p)
{
    if(!p) return;
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        auto it = g_poolLive.find(p);
        if(it != g_poolLive.end())
        {
            const auto key = it->second;
            g_poolLive.erase(it);
            if(g_poolCached + key.second <= g_poolCap)
            {
                g_poolFree[key].push_back(p);
                g_poolCached += key.second;
                return;
            }
        }
    }
    (void)hipFree(p);
}

void setPoolCap(size_t bytes)
{
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lock(g_poolMutex);
        g_poolCap = bytes;
        for(auto& kv : g_poolFree)
            while(g_poolCached > g_poolCap && !kv.second.empty())
            {
                drop.push_back(kv.second.back());
                kv.second.pop_back();
                g_poolCached -= kv.first.second;
            }
    }
    for(void* q : drop) (void)hipFree(q);
}

size_t poolCachedBytes()
{
    std::lock_guard<std::mutex> lock(g_poolMutex);
    return g_poolCached;
}

template<typename X> static hipError_t devAlloc(X** p, size_t count)
{
    return poolAlloc((void**)p, (count ? count : 1) * sizeof(X));
}

// Allocation size of devAlloc (the accounting of residentBytes).
template<typename X> static size_t allocBytes(size_t count)
{
    return sizeClass((count ? count : 1) * sizeof(X));
}

// ---------------------------------------------------------------------------
// Mesh handles.
// ---------------------------------------------------------------------------
// Constraint ids must index the table (or be -1).
static int checkConstraintIds(int64_t nverts, const int32_t* cons, const std::vector<double>* table)
{
    if(!cons) return GM_OK;
    const int64_t n = table ? (int64_t)(table->size() / gmd::CONS_STRIDE) : 0;
    for(int64_t v = 0; v < nverts; ++v)
        if(cons[v] < -1 || cons[v] >= n)
            return fail(GM_ERR_ARG, "vertex " + std::to_string(v) + " names constraint " + std::to_string(cons[v]) +
                                        " of a table of " + std::to_string(n));
    return GM_OK;
}

// Element kind of a whole mesh: a launch of one kind runs the kernels
// specialised for it; prisms (and any mix) take the mixed kernels.
static int meshKind(const gmt::Topology& t)
{
    if(t.nhexs == 0 && t.npris == 0) return gmd::EK_TET;
    if(t.ntets == 0 && t.npris == 0) return gmd::EK_HEX;
    return gmd::EK_MIXED;
}

int makeMesh(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
             int64_t nhexs, const uint32_t* hexs, gm_mesh** out, const int32_t* cons,
             const std::vector<double>* consTable, int64_t npris, const uint32_t* pris)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(nverts < 0 || ntets < 0 || nhexs < 0 || npris < 0) return fail(GM_ERR_ARG, "negative size");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs) || (npris && !pris))
        return fail(GM_ERR_ARG, "null array");
    if(int rc = checkConstraintIds(nverts, cons, consTable)) return rc;
    gm_mesh* m = new gm_mesh();
    if(consTable) m->consTable = *consTable;
    std::string err;
    m->topo = gmt::buildShared(nverts, pos, fixed, cons, ntets, tets, nhexs, hexs, npris, pris, m->hostPos, err);
    if(!m->topo)
    {
        delete m;
        return fail(GM_ERR_ARG, err);
    }
    for(int32_t p : m->topo->groupMaxPatch) m->maxPatch = std::max(m->maxPatch, (int)p);
    m->kind = meshKind(*m->topo);
    m->crew.scaling = m->crew.scaling2 = m->crew.scaling3 = 1.0;
    setAnalyticField(m->crew, m->aspect);
    *out = m;
    return GM_OK;
}

// glm::rotate(m, angle, axis) on the 3x3 block (column-major, m[3 c + r]):
// m times the axis-angle rotation, in the library's published formula.
static void rotate3(double m[9], double angle, double ax, double ay, double az)
{
    const double c = std::cos(angle), s = std::sin(angle);
    const double inv = 1.0 / std::sqrt(ax * ax + ay * ay + az * az);
    const double axis[3] = {ax * inv, ay * inv, az * inv};
    const double temp[3] = {(1.0 - c) * axis[0], (1.0 - c) * axis[1], (1.0 - c) * axis[2]};
    double R[3][3];
    R[0][0] = c + temp[0] * axis[0];
    R[0][1] = temp[0] * axis[1] + s * axis[2];
    R[0][2] = temp[0] * axis[2] - s * axis[1];
    R[1][0] = temp[1] * axis[0] - s * axis[2];
    R[1][1] = c + temp[1] * axis[1];
    R[1][2] = temp[1] * axis[2] + s * axis[0];
    R[2][0] = temp[2] * axis[0] + s * axis[1];
    R[2][1] = temp[2] * axis[1] - s * axis[0];
    R[2][2] = c + temp[2] * axis[2];
    double out[9];
    for(int i = 0; i < 3; ++i)
        for(int r = 0; r < 3; ++r)
            out[3 * i + r] = m[0 * 3 + r] * R[i][0] + m[1 * 3 + r] * R[i][1] + m[2 * 3 + r] * R[i][2];
    std::memcpy(m, out, sizeof(out));
}

void setAnalyticField(gmq::Crew& c, double aspect)
{
    // The rotation starts from the identity, turns by pi/4 about z, then by
    // pi/4 about -y; the inverse is its transpose.
    const double pi = 3.14159265358979323846;
    double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    rotate3(r, pi / 4.0, 0.0, 0.0, 1.0);
    rotate3(r, pi / 4.0, 0.0, -1.0, 0.0);
    c.aspect = aspect;
    for(int j = 0; j < 3; ++j)
        for(int i = 0; i < 3; ++i)
        {
            c.rot[3 * j + i] = r[3 * j + i];
            c.rotInv[3 * j + i] = r[3 * i + j];
        }
}

uint64_t texelHash(const float* texels, size_t n)
{
    uint64_t h = 0xcbf29ce484222325ull;
    const unsigned char* b = (const unsigned char*)texels;
    for(size_t i = 0; i < n * sizeof(float); ++i)
    {
        h ^= b[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

int makeMeshGrouped(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets, const uint32_t* tets,
                    int64_t nhexs, const uint32_t* hexs, int64_t ngroups, const int64_t* groupOff,
                    const uint32_t* groupNodes, gm_mesh** out, const int32_t* cons,
                    const std::vector<double>* consTable, int64_t npris, const uint32_t* pris)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(nverts < 0 || ntets < 0 || nhexs < 0 || npris < 0) return fail(GM_ERR_ARG, "negative size");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs) || (npris && !pris))
        return fail(GM_ERR_ARG, "null array");
    if(int rc = checkConstraintIds(nverts, cons, consTable)) return rc;
    auto t = std::make_shared<gmt::Topology>();
    std::string err;
    if(!gmt::buildGrouped(nverts, pos, fixed, cons, ntets, tets, nhexs, hexs, npris, pris, ngroups, groupOff,
                          groupNodes, *t, err))
        return fail(GM_ERR_ARG, err);
    gm_mesh* m = new gm_mesh();
    if(consTable) m->consTable = *consTable;
    m->hostPos.swap(t->pos);
    m->topo = t;
    for(int32_t p : m->topo->groupMaxPatch) m->maxPatch = std::max(m->maxPatch, (int)p);
    m->kind = meshKind(*m->topo);
    m->crew.scaling = m->crew.scaling2 = m->crew.scaling3 = 1.0;
    setAnalyticField(m->crew, m->aspect);
    *out = m;
    return GM_OK;
}

// Entries of the GETMe corner-slot table: 4 per tetrahedron, 8 per record.
static size_t cornerSlotCount(const gmt::Topology& t)
{
    return std::max<size_t>(4 * (size_t)t.ntets + 8 * (size_t)(t.nhexs + t.npris), 1);
}

size_t residentBytes(const gm_mesh* m, int records, bool getme)
{
    const gmt::Topology& t = *m->topo;
    const size_t V = (size_t)t.nverts;
    const uint64_t elemRecords = (uint64_t)t.nhexs + (uint64_t)t.npris;
    const uint32_t eb = gmd::evalBlocksFor((uint64_t)t.ntets + elemRecords);
    size_t b = allocBytes<double4>(V) + allocBytes<float4>(V) + allocBytes<uint4>((size_t)t.ntets) +
               allocBytes<uint4>(2 * (size_t)elemRecords) + 2 * allocBytes<uint32_t>(V + 1) +
               allocBytes<uint32_t>(t.eref.size()) + allocBytes<uint32_t>(t.nbr.size()) + allocBytes<uint8_t>(V) +
               2 * allocBytes<int32_t>(V) + allocBytes<uint32_t>(V) +
               allocBytes<double>(3 * (size_t)eb) + allocBytes<int32_t>(eb) + allocBytes<int32_t>(20) +
               allocBytes<int32_t>(2);
    if(records > 0) b += allocBytes<double>(3 * (size_t)records) + allocBytes<int32_t>(20 * (size_t)records);
    if(!t.cons.empty()) b += allocBytes<int32_t>(V) + allocBytes<double>(m->consTable.size());
    if(getme)
        b += sizeClass(std::max<size_t>(32, t.eref.size() * sizeof(double4))) +
             sizeClass(std::max<size_t>(32, cornerSlotCount(t) * sizeof(uint32_t))) +
             sizeClass(std::max<size_t>(32, ((size_t)t.ntets + (size_t)elemRecords) * sizeof(double)));
    return b;
}

// Frees the mesh's own device buffers and drops its share of the topology's
// device arrays (dm's tets, hexs, eoff, eref, noff, nbr, fixed, cons,
// cornerSlot and dSlotOf point into them).
static void freeMeshBuffers(gm_mesh* m)
{
    gmd::DevMesh& d = m->dm;
    void* ptrs[] = {d.pos64, d.pos32, d.evalCounts, d.getmePull, d.getmeQual, d.evalPartials, d.evalFlags,
                    d.evalHist, d.evalResult, d.evalHistOut, d.passSlot, (void*)d.consTable,
                    m->elemQualBuf, m->texels, m->dCountsRef};
    for(void* p : ptrs)
        poolFree(p);
    m->devTopo.reset();
    d = gmd::DevMesh{};
    m->elemQualBuf = nullptr;
    m->texels = nullptr;
    m->dSlotOf = nullptr;
    m->dCountsRef = nullptr;
    m->resultCap = 0;
    m->getmeReady = false;
}

void destroyMesh(gm_mesh* m)
{
    if(!m) return;
    if(m->onDevice)
    {
        (void)hipSetDevice(m->device);
        if(m->stream) (void)hipStreamSynchronize(m->stream);
        freePlan(m->plan);
        m->plan = nullptr;
        freeMeshBuffers(m);
        if(m->stream) (void)hipStreamDestroy(m->stream);
        m->stream = nullptr;
        m->onDevice = false;
    }
    delete m;
}

// Device slots and the arrays laid out by them (see gm_runtime.h): the fixed
// vertices, then colour group by colour group, each segment in input order.
// The patch's element order and the neighbour order (the order of the
// reference's sums) are kept.  Computed once per topology and kept with it,
// so meshes of a cached topology upload without it.  nullptr if the groups
// do not cover the movable vertices.
static std::shared_ptr<const DeviceLayout> buildLayout(const gmt::Topology& t)
{
    const size_t V = (size_t)t.nverts;
    auto L = std::make_shared<DeviceLayout>();
    std::vector<uint32_t>& vertexAt = L->vertexAt;
    vertexAt.assign(V, 0);
    const size_t G = t.groupOff.empty() ? 0 : t.groupOff.size() - 1;
    if((size_t)t.nfixed + (G ? (size_t)t.groupOff[G] : 0) != V)
        return nullptr;
    // Segment of each vertex: 0 for the fixed ones, 1 + 3 g + r for run r
    // (boundary, subsurface, interior) of group g, so a group's slots hold
    // its boundary vertices first and every run is contiguous.
    std::vector<uint32_t> segOf(V, 0);
    for(size_t g = 0; g < G; ++g)
        for(int r = 0; r < 3; ++r)
            for(int64_t i = t.runOff[3 * g + r]; i < t.runOff[3 * g + r + 1]; ++i)
                segOf[t.groupNodes[i]] = (uint32_t)(1 + 3 * g + r);
    std::vector<size_t> next(3 * G + 1);
    next[0] = 0;
    for(size_t k = 0; k < 3 * G; ++k) next[k + 1] = (size_t)t.nfixed + (size_t)t.runOff[k];
    for(size_t o = 0; o < V; ++o)
    {
        const size_t v = (size_t)t.permutation[o];
        vertexAt[next[segOf[v]]++] = (uint32_t)v;
    }
    std::vector<uint32_t>& slot = L->slotOf;
    slot.assign(V, 0);
    for(size_t s = 0; s < V; ++s) slot[vertexAt[s]] = (uint32_t)s;
    L->eoff.assign(V + 1, 0);
    L->noff.assign(V + 1, 0);
    L->eref.resize(t.eref.size());
    L->nbr.resize(t.nbr.size());
    L->fixed.resize(V);
    L->cornerSlot.assign(cornerSlotCount(t), 0);
    for(size_t s = 0; s < V; ++s)
    {
        const size_t v = vertexAt[s];
        L->fixed[s] = t.fixed[v];
        if(!t.cons.empty())
        {
            if(L->cons.empty()) L->cons.assign(V, -1);
            L->cons[s] = t.cons[v];
        }
        const uint32_t e0 = t.eoff[v], e1 = t.eoff[v + 1];
        L->eoff[s + 1] = L->eoff[s] + (e1 - e0);
        for(uint32_t r = e0; r < e1; ++r)
        {
            const uint32_t at = L->eoff[s] + (r - e0);
            const uint32_t ref = t.eref[r];
            L->eref[at] = ref;
            const size_t id = ref >> 4, k = ref & 7;
            // The patch entry of each element corner (GETMe's element step
            // writes there): tets 4 e + k, then records (hexahedra, then
            // prisms) 4 T + 8 r + k.
            L->cornerSlot[(ref & 8) ? 4 * (size_t)t.ntets + 8 * id + k : 4 * id + k] = at;
        }
        L->noff[s + 1] = L->noff[s] + (t.noff[v + 1] - t.noff[v]);
        for(uint32_t r = t.noff[v]; r < t.noff[v + 1]; ++r) L->nbr[L->noff[s] + (r - t.noff[v])] = slot[t.nbr[r]];
    }
    L->tets.resize(t.tets.size());
    L->hexs.resize(t.hexs.size());
    for(size_t k = 0; k < t.tets.size(); ++k) L->tets[k] = slot[t.tets[k]];
    for(size_t k = 0; k < t.hexs.size(); ++k) L->hexs[k] = slot[t.hexs[k]];
    return L;
}

static std::mutex g_layoutMu;

// Device copies held by cached topologies are released at exit by a handler
// registered after the HIP runtime has initialised, so it runs before the
// runtime's own teardown; the cache's static destructor then frees no device
// memory.
static void releaseCachedDeviceCopies()
{
    gmt::setCacheCapacity(0);
}

// The device arrays of layout L (of topology t) on the current device:
// found, or uploaded on stream s and published once complete.  Uploads of one
// topology are serialised by L.devMu, so concurrent first uploads make one
// copy.
static hipError_t deviceTopoOf(const DeviceLayout& L, const gmt::Topology& t, int device, hipStream_t s,
                               std::shared_ptr<DeviceTopo>& out)
{
    static std::once_flag atExit;
    std::call_once(atExit, [] { std::atexit(releaseCachedDeviceCopies); });
    std::lock_guard<std::mutex> lock(L.devMu);
    auto it = L.onDevice.find(device);
    if(it != L.onDevice.end())
    {
        out = it->second;
        return hipSuccess;
    }
    auto D = std::make_shared<DeviceTopo>();
    const size_t V = (size_t)t.nverts;
    hipError_t e = hipSuccess;
    auto put = [&](auto** dst, const auto& src, size_t count) {
        if(e != hipSuccess) return;
        e = devAlloc(dst, count);
        if(e == hipSuccess && count)
            e = hipMemcpyAsync(*dst, src.data(), count * sizeof(**dst), hipMemcpyHostToDevice, s);
    };
    static_assert(sizeof(uint4) == 16, "uint4 is 4 ids");
    e = devAlloc(&D->tets, (size_t)t.ntets);
    if(e == hipSuccess && t.ntets)
        e = hipMemcpyAsync(D->tets, L.tets.data(), (size_t)t.ntets * 16, hipMemcpyHostToDevice, s);
    const size_t records = (size_t)(t.nhexs + t.npris);
    if(e == hipSuccess) e = devAlloc(&D->hexs, 2 * records);
    if(e == hipSuccess && records)
        e = hipMemcpyAsync(D->hexs, L.hexs.data(), records * 32, hipMemcpyHostToDevice, s);
    put(&D->eoff, L.eoff, V + 1);
    put(&D->eref, L.eref, L.eref.size());
    put(&D->noff, L.noff, V + 1);
    put(&D->nbr, L.nbr, L.nbr.size());
    put(&D->fixed, L.fixed, V);
    put(&D->slotOf, L.slotOf, V);
    if(!L.cons.empty()) put(&D->cons, L.cons, V);
    if(e == hipSuccess) e = hipStreamSynchronize(s);
    if(e != hipSuccess) return e;
    L.onDevice[device] = D;
    out = D;
    return hipSuccess;
}

// The GETMe corner slots of a topology's device copy (uploaded once).
static hipError_t cornerSlotsOf(const DeviceLayout& L, DeviceTopo& D, hipStream_t s)
{
    std::lock_guard<std::mutex> lock(L.devMu);
    if(D.cornerSlot) return hipSuccess;
    uint32_t* p = nullptr;
    hipError_t e = devAlloc(&p, L.cornerSlot.size());
    if(e == hipSuccess)
        e = hipMemcpyAsync(p, L.cornerSlot.data(), L.cornerSlot.size() * 4, hipMemcpyHostToDevice, s);
    if(e == hipSuccess) e = hipStreamSynchronize(s);
    if(e != hipSuccess)
    {
        poolFree(p);
        return e;
    }
    D.cornerSlot = p;
    return hipSuccess;
}

static std::shared_ptr<const DeviceLayout> layoutOf(const gmt::Topology& t)
{
    {
        std::lock_guard<std::mutex> lock(g_layoutMu);
        if(t.deviceLayout) return std::static_pointer_cast<const DeviceLayout>(t.deviceLayout);
    }
    std::shared_ptr<const DeviceLayout> L = buildLayout(t);
    if(!L) return nullptr;
    std::lock_guard<std::mutex> lock(g_layoutMu);
    if(t.deviceLayout) return std::static_pointer_cast<const DeviceLayout>(t.deviceLayout);
    t.deviceLayout = L;
    return L;
}

int prepareDeviceTopology(const gm_mesh* m)
{
    const gmt::Topology& t = *m->topo;
    std::shared_ptr<const DeviceLayout> L = layoutOf(t);
    if(!L) return fail(GM_ERR_ARG, "colour groups do not cover the movable vertices");
    int dev = 0;
    hipStream_t s = nullptr;
    hipError_t e = hipGetDevice(&dev);
    if(e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    std::shared_ptr<DeviceTopo> D;
    if(e == hipSuccess) e = deviceTopoOf(*L, t, dev, s, D);
    if(s) (void)hipStreamDestroy(s);
    if(e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP,
                    std::string("topology upload: ") + hipGetErrorString(e));
    return GM_OK;
}

int uploadMesh(gm_mesh* m)
{
    const gmt::Topology& t = *m->topo;
    for(int64_t v = 0; v < t.nverts; ++v)
        if(!t.fixed[v] && t.eoff[v + 1] == t.eoff[v])
            return fail(GM_ERR_ARG, "a free vertex belongs to no element");
    if((uint64_t)t.nverts >= (1ull << 32)) return fail(GM_ERR_ARG, "too many vertices for 32-bit ids");
    if(m->onDevice) return fail(GM_ERR_ARG, "mesh is already resident");
    if(m->hostPos.size() != 3 * (size_t)t.nverts) return fail(GM_ERR_ARG, "mesh has no initial positions");

    m->onDevice = true;
    auto bail = [&](int code) {
        if(m->stream) (void)hipStreamSynchronize(m->stream);
        freePlan(m->plan);
        m->plan = nullptr;
        freeMeshBuffers(m);
        if(m->stream) (void)hipStreamDestroy(m->stream);
        m->stream = nullptr;
        m->onDevice = false;
        return code;
    };
#define GM_TRY(call)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if(e_ != hipSuccess)                                                                            \
            return bail(fail(e_ == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP,                     \
                             std::string(#call) + ": " + hipGetErrorString(e_)));                       \
    } while(0)
    GM_TRY(hipGetDevice(&m->device));
    GM_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));

    const size_t V = (size_t)t.nverts;
    gmd::DevMesh& d = m->dm;
    d = gmd::DevMesh{};
    d.nverts = (uint32_t)t.nverts;
    d.ntets = (uint32_t)t.ntets;
    d.nhexs = (uint32_t)t.nhexs;
    d.npris = (uint32_t)t.npris;
    d.evalBlocks = gmd::evalBlocksFor((uint64_t)t.ntets + (uint64_t)t.nhexs + (uint64_t)t.npris);
    std::shared_ptr<const DeviceLayout> L = layoutOf(t);
    if(!L) return bail(fail(GM_ERR_ARG, "colour groups do not cover the movable vertices"));
    m->layout = L;
    m->slotOf = L->slotOf;
    std::vector<double> p64(4 * V);
    std::vector<float> p32(4 * V);
    for(size_t s_ = 0; s_ < V; ++s_)
    {
        const size_t v = L->vertexAt[s_];
        for(int k = 0; k < 3; ++k)
        {
            p64[4 * s_ + k] = m->hostPos[3 * v + k];
            p32[4 * s_ + k] = (float)m->hostPos[3 * v + k];
        }
        p64[4 * s_ + 3] = 0.0;
        p32[4 * s_ + 3] = 0.0f;
    }
    // The topology's arrays: shared with every mesh of this topology on this
    // device, uploaded by the first of them.
    GM_TRY(deviceTopoOf(*L, t, m->device, m->stream, m->devTopo));
    const DeviceTopo& D = *m->devTopo;
    d.tets = D.tets;
    d.hexs = D.hexs;
    d.eoff = D.eoff;
    d.eref = D.eref;
    d.noff = D.noff;
    d.nbr = D.nbr;
    d.fixed = D.fixed;
    d.cons = D.cons;
    m->dSlotOf = D.slotOf;
    GM_TRY(devAlloc(&d.pos64, V));
    GM_TRY(devAlloc(&d.pos32, V));
    double* ct = nullptr;
    if(!L->cons.empty())
    {
        GM_TRY(devAlloc(&ct, m->consTable.size()));
        d.consTable = ct;
    }
    GM_TRY(devAlloc(&d.evalCounts, V));
    GM_TRY(devAlloc(&m->dCountsRef, V));
    GM_TRY(devAlloc(&d.evalPartials, 3 * (size_t)d.evalBlocks));
    GM_TRY(devAlloc(&d.evalFlags, (size_t)d.evalBlocks));
    GM_TRY(devAlloc(&d.evalHist, 20));
    GM_TRY(devAlloc(&d.passSlot, 2));

    hipStream_t s = m->stream;
    if(V)
    {
        GM_TRY(hipMemcpyAsync(d.pos64, p64.data(), V * sizeof(double4), hipMemcpyHostToDevice, s));
        GM_TRY(hipMemcpyAsync(d.pos32, p32.data(), V * sizeof(float4), hipMemcpyHostToDevice, s));
    }
    if(ct)
        GM_TRY(hipMemcpyAsync(ct, m->consTable.data(), m->consTable.size() * sizeof(double), hipMemcpyHostToDevice, s));
    GM_TRY(hipStreamSynchronize(s));
#undef GM_TRY
    std::vector<gm_mesh*> one{m};
    int rc = buildPlan(one, m->stream, &m->plan);
    if(rc) return bail(rc);
    m->hostPos.clear();
    m->hostPos.shrink_to_fit();
    return GM_OK;
}

// Patch-size boundaries at which the vertex kernels change shape: wave
// kernel lanes per vertex (8, 16, 32, 64), gradient-descent tile slots
// (8, 16, 32) and spawn-search tiles (64, 128, 256).
static const int PATCH_CLASSES = 7;
static int patchClass(int maxPatch)
{
    const int bounds[PATCH_CLASSES - 1] = {8, 16, 32, 64, 128, 256};
    int c = 0;
    while(c < PATCH_CLASSES - 1 && maxPatch > bounds[c]) ++c;
    return c;
}

static uint32_t ceilDiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Appends one block span per member with blocks; returns the work description
// (span pointer filled in after upload).
template<typename F>
static gmd::BlockWork blockSpans(const std::vector<gm_mesh*>& meshes, std::vector<uint4>& spans, size_t& first,
                                 F blocksOf)
{
    first = spans.size();
    uint32_t total = 0;
    for(size_t i = 0; i < meshes.size(); ++i)
    {
        const uint32_t nb = blocksOf(meshes[i]);
        if(nb == 0) continue;
        spans.push_back(make_uint4((uint32_t)i, 0, total, 0));
        total += nb;
    }
    gmd::BlockWork w{};
    w.nspans = (uint32_t)(spans.size() - first);
    w.count = total;
    return w;
}

// A run of consecutive device slots of one member: a colour group's
// boundary vertices, or its other movable vertices, or (GETMe with sliding
// vertices) one type of one group.
struct SlotRun
{
    uint32_t first;
    uint64_t count;
    int maxPatch;
    int constrained;
};

// Appends the work lists of one step (an independent set of every member)
// to `out`: one list per (sliding or not, element kind, patch class).  The
// kind keeps the candidate-parallel kernels on patches of a single kind; the
// patch class (largest patch of the member's run, rounded up to the kernels'
// tile boundaries) makes each list pick, for every member, the kernel that
// member picks when it runs alone.  Within a list, vertices keep the member
// order and each run's slot order.
template<class F>
static int stepLists(const std::vector<gm_mesh*>& meshes, std::vector<uint4>& spans, std::vector<VList>& out,
                     F runsOf)
{
    std::vector<std::vector<SlotRun>> runs(meshes.size());
    for(size_t i = 0; i < meshes.size(); ++i) runs[i] = runsOf(meshes[i]);
    for(int constrained = 1; constrained >= 0; --constrained)
        for(int kind = gmd::EK_TET; kind <= gmd::EK_MIXED; ++kind)
            for(int cls = 0; cls < PATCH_CLASSES; ++cls)
            {
                VList l;
                l.kind = kind;
                l.constrained = constrained;
                l.firstSpan = spans.size();
                for(size_t i = 0; i < meshes.size(); ++i)
                {
                    if(meshes[i]->kind != kind) continue;
                    for(const SlotRun& r : runs[i])
                    {
                        if(r.count == 0 || r.constrained != constrained || patchClass(r.maxPatch) != cls) continue;
                        // Each run starts at a multiple of ITEM_ALIGN (the
                        // items of a wavefront never span two runs).
                        const uint64_t z = ((uint64_t)l.count + gmd::ITEM_ALIGN - 1) / gmd::ITEM_ALIGN * gmd::ITEM_ALIGN;
                        if(z + r.count >= (1ull << 32)) return fail(GM_ERR_ARG, "colour step too large");
                        spans.push_back(make_uint4((uint32_t)i, r.first, (uint32_t)z, (uint32_t)r.count));
                        l.count = (uint32_t)(z + r.count);
                        l.maxPatch = std::max(l.maxPatch, r.maxPatch);
                    }
                }
                l.nsegs = (uint32_t)(spans.size() - l.firstSpan);
                if(l.count) out.push_back(l);
            }
    return GM_OK;
}

static int buildPlanInto(Plan* p, const std::vector<gm_mesh*>& meshes)
{
    int64_t maxColours = 0;
    bool sliding = false;
    for(gm_mesh* m : meshes)
    {
        maxColours = std::max(maxColours, m->topo->ngroups());
        sliding = sliding || !m->topo->cons.empty();
    }
    p->colours.resize((size_t)maxColours);
    std::vector<uint4> spans;
    // Vertex-wise sweeps: colour by colour, each group's boundary vertices in
    // lists of their own (the generic kernels, which snap every trial
    // position onto the vertex's constraint), the rest in the others.
    for(int64_t g = 0; g < maxColours; ++g)
    {
        int rc = stepLists(meshes, spans, p->colours[g], [g](const gm_mesh* m) {
            const gmt::Topology& t = *m->topo;
            std::vector<SlotRun> r;
            if(g >= t.ngroups()) return r;
            const int64_t* o = &t.runOff[3 * g];
            r.push_back(SlotRun{(uint32_t)(t.nfixed + o[0]), (uint64_t)(o[1] - o[0]), t.runMaxPatch[3 * g], 1});
            r.push_back(SlotRun{(uint32_t)(t.nfixed + o[1]), (uint64_t)(o[3] - o[1]),
                                std::max(t.runMaxPatch[3 * g + 1], t.runMaxPatch[3 * g + 2]), 0});
            return r;
        });
        if(rc) return rc;
    }
    // GETMe's vertex steps follow NodeGroups' serial order (its reference
    // implementation is serial): every group's boundary run, then every
    // subsurface run, then every interior run.  Without sliding vertices
    // every movable vertex is interior and that order is the colour order.
    if(sliding)
        for(int type = 0; type < 3; ++type)
            for(int64_t g = 0; g < maxColours; ++g)
            {
                std::vector<VList> step;
                int rc = stepLists(meshes, spans, step, [g, type](const gm_mesh* m) {
                    const gmt::Topology& t = *m->topo;
                    std::vector<SlotRun> r;
                    if(g >= t.ngroups()) return r;
                    const int64_t k = 3 * g + type;
                    r.push_back(SlotRun{(uint32_t)(t.nfixed + t.runOff[k]), (uint64_t)(t.runOff[k + 1] - t.runOff[k]),
                                        t.runMaxPatch[k], type == 0 ? 1 : 0});
                    return r;
                });
                if(rc) return rc;
                if(!step.empty()) p->getmeSteps.push_back(std::move(step));
            }
    size_t fTet, fHex, fPri, fEval, fVert;
    p->tetBlocks = blockSpans(meshes, spans, fTet, [](const gm_mesh* m) { return ceilDiv(m->topo->ntets, gmd::ELEM_BLOCK); });
    p->hexBlocks = blockSpans(meshes, spans, fHex, [](const gm_mesh* m) { return ceilDiv(m->topo->nhexs, gmd::ELEM_BLOCK); });
    p->priBlocks = blockSpans(meshes, spans, fPri, [](const gm_mesh* m) { return ceilDiv(m->topo->npris, gmd::ELEM_BLOCK); });
    p->evalBlocks = blockSpans(meshes, spans, fEval, [](const gm_mesh* m) { return m->dm.evalBlocks; });
    p->vertBlocks = blockSpans(meshes, spans, fVert, [](const gm_mesh* m) { return ceilDiv(m->topo->nverts, gmd::VERT_BLOCK); });
    std::vector<uint32_t> ml(meshes.size());
    for(size_t i = 0; i < meshes.size(); ++i) ml[i] = (uint32_t)i;

    GMR_HIP(devAlloc(&p->spans, spans.size()));
    GMR_HIP(devAlloc(&p->meshList, ml.size()));
    p->nMeshes = (uint32_t)ml.size();
    if(!spans.empty())
        GMR_HIP(hipMemcpyAsync(p->spans, spans.data(), spans.size() * sizeof(uint4), hipMemcpyHostToDevice, p->stream));
    GMR_HIP(hipMemcpyAsync(p->meshList, ml.data(), ml.size() * sizeof(uint32_t), hipMemcpyHostToDevice, p->stream));
    for(auto& col : p->colours)
        for(VList& l : col) l.segs = p->spans + l.firstSpan;
    for(auto& step : p->getmeSteps)
        for(VList& l : step) l.segs = p->spans + l.firstSpan;
    p->tetBlocks.spans = p->spans + fTet;
    p->hexBlocks.spans = p->spans + fHex;
    p->priBlocks.spans = p->spans + fPri;
    p->evalBlocks.spans = p->spans + fEval;
    p->vertBlocks.spans = p->spans + fVert;
    GMR_HIP(devAlloc(&p->dTable, meshes.size()));
    GMR_HIP(hipEventCreate(&p->ev[0]));
    GMR_HIP(hipEventCreate(&p->ev[1]));
    GMR_HIP(hipStreamSynchronize(p->stream));
    return GM_OK;
}

int buildPlan(const std::vector<gm_mesh*>& meshes, hipStream_t stream, Plan** out)
{
    *out = nullptr;
    if(meshes.empty()) return fail(GM_ERR_ARG, "a plan needs at least one mesh");
    for(const gm_mesh* m : meshes)
        if((uint64_t)m->topo->nverts >= (1ull << 32) || (uint64_t)m->topo->ntets >= (1ull << 32) ||
           (uint64_t)m->topo->nhexs + (uint64_t)m->topo->npris >= (1ull << 32))
            return fail(GM_ERR_ARG, "mesh too large for 32-bit work items");
    if(meshes.size() >= (1ull << 32)) return fail(GM_ERR_ARG, "too many meshes");
    Plan* p = new Plan();
    p->stream = stream;
    p->meshes = meshes;
    p->device = meshes[0]->device;
    int rc = buildPlanInto(p, meshes);
    if(rc)
    {
        freePlan(p);
        return rc;
    }
    *out = p;
    return GM_OK;
}

void clearGraphs(Plan* p)
{
    for(auto& kv : p->graphs) (void)hipGraphExecDestroy(kv.second);
    for(hipGraph_t g : p->graphDefs) (void)hipGraphDestroy(g);
    p->graphs.clear();
    p->graphDefs.clear();
}

void freePlan(Plan* p)
{
    if(!p) return;
    clearGraphs(p);
    for(void* q : p->owned) poolFree(q);
    poolFree(p->dTable);
    poolFree(p->spans);
    poolFree(p->meshList);
    for(hipEvent_t e : p->ev) if(e) (void)hipEventDestroy(e);
    for(hipEvent_t e : p->pool) if(e) (void)hipEventDestroy(e);
    delete p;
}

static int ensureRecords(gm_mesh* m, int records)
{
    if(records <= m->resultCap) return GM_OK;
    poolFree(m->dm.evalResult);
    poolFree(m->dm.evalHistOut);
    m->dm.evalResult = nullptr;
    m->dm.evalHistOut = nullptr;
    m->resultCap = 0;
    m->dm.resultCap = 0;
    GMR_HIP(devAlloc(&m->dm.evalResult, 3 * (size_t)records));
    GMR_HIP(devAlloc(&m->dm.evalHistOut, 20 * (size_t)records));
    m->resultCap = records;
    m->dm.resultCap = (uint32_t)records;
    return GM_OK;
}

static int ensureGetme(gm_mesh* m)
{
    if(m->getmeReady) return GM_OK;
    // (vpp, weight) per patch entry, sized for FP64 (FP32 runs use the front
    // half), and the patch entry of every element corner: the element step
    // writes each corner's result where the vertex step reads it, so a
    // vertex pulls one contiguous run.
    const gmt::Topology& t = *m->topo;
    const size_t N = t.eref.size();
    const size_t pullBytes = N * sizeof(double4);
    const size_t qualBytes = ((size_t)t.ntets + (size_t)t.nhexs + (size_t)t.npris) * sizeof(double);
    GMR_HIP(cornerSlotsOf(*m->layout, *m->devTopo, m->stream));
    m->dm.cornerSlot = m->devTopo->cornerSlot;
    GMR_HIP(poolAlloc(&m->dm.getmePull, std::max<size_t>(32, pullBytes)));
    GMR_HIP(poolAlloc(&m->dm.getmeQual, std::max<size_t>(32, qualBytes)));
    m->getmeReady = true;
    return GM_OK;
}

int preparePlan(Plan* p, int records, bool getme)
{
    for(gm_mesh* m : p->meshes)
    {
        int rc = ensureRecords(m, records);
        if(rc) return rc;
        if(getme && (rc = ensureGetme(m))) return rc;
    }
    p->table.resize(p->meshes.size());
    for(size_t i = 0; i < p->meshes.size(); ++i) p->table[i] = p->meshes[i]->dm;
    GMR_HIP(hipMemcpyAsync(p->dTable, p->table.data(), p->table.size() * sizeof(gmd::DevMesh),
                           hipMemcpyHostToDevice, p->stream));
    GMR_HIP(hipStreamSynchronize(p->stream));
    return GM_OK;
}

// ----- per-launch timing (eager launches only)
static int timedBegin(Plan* p, bool on, int tag)
{
    if(!on) return -1;
    if(p->poolUsed + 2 > p->pool.size())
    {
        size_t old = p->pool.size();
        p->pool.resize(old + 256, nullptr);
        p->poolTag.resize(old + 256, -1);
        for(size_t i = old; i < p->pool.size(); ++i)
            if(hipEventCreate(&p->pool[i]) != hipSuccess)
            {
                p->pool.resize(i);
                p->poolTag.resize(i);
                break;
            }
        if(p->poolUsed + 2 > p->pool.size()) return -1;
    }
    int slot = (int)p->poolUsed;
    p->poolUsed += 2;
    p->poolTag[slot] = tag;
    (void)hipEventRecord(p->pool[slot], p->stream);
    return slot;
}

static void timedEnd(Plan* p, int slot)
{
    if(slot >= 0) (void)hipEventRecord(p->pool[slot + 1], p->stream);
}

static void collectTimings(Plan* p)
{
    for(int k = 0; k < GM_KT_COUNT; ++k)
    {
        p->kernMs[k] = 0.0;
        p->kernLaunches[k] = 0;
    }
    for(size_t i = 0; i + 1 < p->poolUsed; i += 2)
    {
        float ms = 0.0f;
        if(p->poolTag[i] >= 0 && hipEventElapsedTime(&ms, p->pool[i], p->pool[i + 1]) == hipSuccess)
        {
            p->kernMs[p->poolTag[i]] += ms;
            p->kernLaunches[p->poolTag[i]] += 1;
        }
    }
    p->poolUsed = 0;
}

bool planHasRecords(const Plan* p)
{
    for(const gm_mesh* m : p->meshes)
        if(m->topo->nhexs + m->topo->npris > 0) return true;
    return false;
}

static hipError_t evaluateAll(Plan* p, const gmq::Crew& c, int ev, int sa, bool fromF32, bool timed)
{
    int slot = timedBegin(p, timed, GM_KT_EVALUATE);
    const bool records = planHasRecords(p);
    hipError_t e = fromF32
        ? gmd::launchEvaluate<float>(p->dTable, p->evalBlocks, p->meshList, p->nMeshes, c, ev, sa, records,
                                     p->stream)
        : gmd::launchEvaluate<double>(p->dTable, p->evalBlocks, p->meshList, p->nMeshes, c, ev, sa, records,
                                      p->stream);
    timedEnd(p, slot);
    return e;
}

// One sweep over colours [g0, g1) of every member (GETMe: the element step
// over all members first, as GetmeSmoother::smoothMeshThread does per mesh).
static hipError_t sweep(Plan* p, const gmq::Crew& c, int ev, int sa, const RunParams& r, bool f64, int64_t g0,
                        int64_t g1, bool timed)
{
    hipError_t e = hipSuccess;
    const bool getme = r.smoother == gmd::SM_GETME;
    if(getme)
    {
        int slot = timedBegin(p, timed, f64 ? GM_KT_GETME_ELEM_F64 : GM_KT_GETME_ELEM_F32);
        e = f64 ? gmd::launchGetmeElements<double>(p->dTable, p->tetBlocks, p->hexBlocks, p->priBlocks, c, ev, sa,
                                                   gmr::passFast(r.precision), p->stream)
                : gmd::launchGetmeElements<float>(p->dTable, p->tetBlocks, p->hexBlocks, p->priBlocks, c, ev, sa,
                                                  false, p->stream);
        timedEnd(p, slot);
    }
    // GETMe always sweeps every group, in serial order when members slide.
    const std::vector<std::vector<VList>>& steps = getme && !p->getmeSteps.empty() ? p->getmeSteps : p->colours;
    if(getme)
    {
        g0 = 0;
        g1 = (int64_t)steps.size();
    }
    for(int64_t g = g0; e == hipSuccess && g < g1; ++g)
    {
        for(const VList& l : steps[g])
        {
            gmd::VertexWork w{p->dTable, l.segs, l.nsegs, l.count, l.kind, l.maxPatch,
                              f64 && gmr::passFast(r.precision) ? 1 : 0, l.constrained};
            gmd::LaneShape lane = r.lane;
            if(lane.tileElems == 0)
            {
                // The tile that just holds the list's largest patch (rounded
                // to 4 elements): spawn search stages one patch per
                // wavefront, gradient descent eight.  Less LDS per wavefront
                // keeps more of them resident.
                const int patch = (l.maxPatch + 3) & ~3;
                lane.tileElems = r.smoother == gmd::SM_GRADIENT_DESCENT ? 8 * patch : patch;
            }
            int tag = getme ? (f64 ? GM_KT_GETME_VERT_F64 : GM_KT_GETME_VERT_F32)
                            : (f64 ? GM_KT_SMOOTH_F64 : GM_KT_SMOOTH_F32);
            int slot = timedBegin(p, timed, tag);
            if(getme)
                e = f64 ? gmd::launchGetmeVertices<double>(w, c, ev, sa, r.kernel, lane, p->stream)
                        : gmd::launchGetmeVertices<float>(w, c, ev, sa, r.kernel, lane, p->stream);
            else
                e = f64 ? gmd::launchSmooth<double>(w, c, ev, sa, r.smoother, r.kernel, lane, r.spawnCoeff,
                                                    p->stream)
                        : gmd::launchSmooth<float>(w, c, ev, sa, r.smoother, r.kernel, lane, r.spawnCoeff,
                                                   p->stream);
            timedEnd(p, slot);
            if(e != hipSuccess) break;
        }
    }
    return e;
}

// One pass body: the sweep (FP64 or FP32) and the FP64 quality reduction of
// its result.  With r.graph the body is captured once per configuration and
// replayed; the reduction's record slot advances on the device, so every
// replay writes the next record.
static int passBody(Plan* p, const gmq::Crew& c, int ev, int sa, const RunParams& r, bool f64, bool evalF32)
{
    const int64_t G = (int64_t)p->colours.size();
    if(!r.graph)
    {
        hipError_t e = sweep(p, c, ev, sa, r, f64, 0, G, r.timeKernels);
        if(e == hipSuccess) e = evaluateAll(p, c, ev, sa, evalF32, r.timeKernels);
        if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("pass launch: ") + hipGetErrorString(e));
        return GM_OK;
    }
    if(!p->graphs.empty() && (p->graphEv != ev || p->graphSa != sa || std::memcmp(&p->graphCrew, &c, sizeof(c)) != 0))
        clearGraphs(p);
    p->graphCrew = c;
    p->graphEv = ev;
    p->graphSa = sa;
    uint64_t coeffBits;
    std::memcpy(&coeffBits, &r.spawnCoeff, sizeof(coeffBits));
    auto key = std::make_tuple(r.smoother, (int)f64, r.kernel, r.lane.patchesPerGroup, r.lane.tileElems, coeffBits,
                               (int)evalF32);
    auto it = p->graphs.find(key);
    if(it == p->graphs.end())
    {
        hipGraph_t graph = nullptr;
        GMR_HIP(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = sweep(p, c, ev, sa, r, f64, 0, G, false);
        if(e == hipSuccess) e = evaluateAll(p, c, ev, sa, evalF32, false);
        hipError_t e2 = hipStreamEndCapture(p->stream, &graph);
        if(e != hipSuccess || e2 != hipSuccess)
        {
            if(graph) (void)hipGraphDestroy(graph);
            return fail(GM_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if(e != hipSuccess)
        {
            (void)hipGraphDestroy(graph);
            return fail(GM_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        p->graphDefs.push_back(graph);
        it = p->graphs.emplace(key, exec).first;
    }
    GMR_HIP(hipGraphLaunch(it->second, p->stream));
    return GM_OK;
}

namespace
{
std::atomic<int> g_polishPasses{1};
}

int polishPasses() { return g_polishPasses.load(); }
void setPolishPasses(int n) { g_polishPasses.store(n); }

int runPlan(Plan* p, const gmq::Crew& c, int ev, int sa, const RunParams& r)
{
    for(gm_mesh* m : p->meshes)
    {
        m->dm.autopilot = r.autopilot ? 1 : 0;
        m->dm.apMeanGain = r.apMeanGain;
        m->dm.apMinGain = r.apMinGain;
    }
    int rc = preparePlan(p, r.passes + 1, r.smoother == gmd::SM_GETME);
    if(rc) return rc;
    hipStream_t s = p->stream;
    const bool timed = r.timeKernels && !r.graph;
    p->poolUsed = 0;
    GMR_HIP(hipEventRecord(p->ev[0], s));
    GMR_HIP(gmd::launchResetRun(p->dTable, p->vertBlocks, p->meshList, p->nMeshes, s));
    bool inF32 = false;
    if(startsF32(r.precision))
    {
        GMR_HIP(gmd::launchToF32(p->dTable, p->vertBlocks, s));
        inF32 = true;
    }
    GMR_HIP(evaluateAll(p, c, ev, sa, inF32, timed));
    for(int pass = 0; pass < r.passes; ++pass)
    {
        const bool f64 = passF64(r.precision, pass, r.passes, r.polishPasses);
        if(f64 && inF32)
        {
            GMR_HIP(gmd::launchToF64(p->dTable, p->vertBlocks, s));
            inF32 = false;
        }
        RunParams rr = r;
        rr.timeKernels = timed;
        if((rc = passBody(p, c, ev, sa, rr, f64, inF32))) return rc;
    }
    if(inF32)
        GMR_HIP(gmd::launchToF64(p->dTable, p->vertBlocks, s));
    GMR_HIP(hipEventRecord(p->ev[1], s));
    GMR_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    GMR_HIP(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
    p->totalMs = ms;
    collectTimings(p);
    return GM_OK;
}

int launchBegin(Plan* p, const gmq::Crew& c, int ev, int sa, bool toF32)
{
    hipStream_t s = p->stream;
    GMR_HIP(gmd::launchResetRun(p->dTable, p->vertBlocks, p->meshList, p->nMeshes, s));
    if(toF32) GMR_HIP(gmd::launchToF32(p->dTable, p->vertBlocks, s));
    GMR_HIP(evaluateAll(p, c, ev, sa, toF32, false));
    return GM_OK;
}

int launchPass(Plan* p, const gmq::Crew& c, int ev, int sa, const RunParams& r, bool f64, bool convertFirst)
{
    if(convertFirst) GMR_HIP(gmd::launchToF64(p->dTable, p->vertBlocks, p->stream));
    RunParams rr = r;
    rr.graph = false;
    rr.timeKernels = false;
    return passBody(p, c, ev, sa, rr, f64, !f64);
}

int launchToF64All(Plan* p)
{
    GMR_HIP(gmd::launchToF64(p->dTable, p->vertBlocks, p->stream));
    return GM_OK;
}

int sweepColours(Plan* p, const gmq::Crew& c, int ev, int sa, const RunParams& r, int64_t g0, int64_t g1)
{
    for(gm_mesh* m : p->meshes) m->dm.autopilot = 0;
    int rc = preparePlan(p, 1, r.smoother == gmd::SM_GETME);
    if(rc) return rc;
    hipStream_t s = p->stream;
    p->poolUsed = 0;
    GMR_HIP(gmd::launchResetRun(p->dTable, p->vertBlocks, p->meshList, p->nMeshes, s));
    const bool f64 = !startsF32(r.precision);
    if(!f64) GMR_HIP(gmd::launchToF32(p->dTable, p->vertBlocks, s));
    hipError_t e = sweep(p, c, ev, sa, r, f64, std::min<int64_t>(g0, (int64_t)p->colours.size()),
                         std::min<int64_t>(g1, (int64_t)p->colours.size()), false);
    if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("sweep: ") + hipGetErrorString(e));
    if(!f64) GMR_HIP(gmd::launchToF64(p->dTable, p->vertBlocks, s));
    GMR_HIP(hipStreamSynchronize(s));
    return GM_OK;
}

int recordsWritten(Plan* p, size_t i, int* records)
{
    int32_t slot[2] = {0, 0};
    GMR_HIP(hipMemcpyAsync(slot, p->meshes[i]->dm.passSlot, sizeof(slot), hipMemcpyDeviceToHost, p->stream));
    GMR_HIP(hipStreamSynchronize(p->stream));
    *records = slot[0];
    return GM_OK;
}

int fetchRecords(Plan* p, size_t i, int records, double* stats, int32_t* hist, int32_t* evalCounts)
{
    gm_mesh* m = p->meshes[i];
    std::vector<double> r(3 * (size_t)records);
    std::vector<int32_t> h(20 * (size_t)records);
    GMR_HIP(hipMemcpyAsync(r.data(), m->dm.evalResult, r.size() * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    GMR_HIP(hipMemcpyAsync(h.data(), m->dm.evalHistOut, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                           p->stream));
    if(evalCounts && m->topo->nverts)
    {
        const uint32_t V = (uint32_t)m->topo->nverts;
        GMR_HIP(gmd::launchGatherCounts(m->dm.evalCounts, m->dSlotOf, m->dCountsRef, V, p->stream));
        GMR_HIP(hipMemcpyAsync(evalCounts, m->dCountsRef, (size_t)V * sizeof(int32_t), hipMemcpyDeviceToHost,
                               p->stream));
    }
    GMR_HIP(hipStreamSynchronize(p->stream));
    for(int k = 0; k < records; ++k)
    {
        if(stats)
        {
            // result record = {min, sum 1/q, n}; harmonic mean = n / sum 1/q.
            stats[3 * k] = r[3 * k];
            stats[3 * k + 1] = r[3 * k + 2] / r[3 * k + 1];
            stats[3 * k + 2] = r[3 * k + 1];
        }
        if(hist) std::memcpy(hist + 20 * k, h.data() + 20 * k, 20 * sizeof(int32_t));
    }
    return GM_OK;
}

} // namespace gmr

DeviceTopo::~DeviceTopo()
{
    void* ptrs[] = {tets, hexs, eoff, eref, noff, nbr, slotOf, fixed, cons, cornerSlot};
    for(void* p : ptrs) gmr::poolFree(p);
}
