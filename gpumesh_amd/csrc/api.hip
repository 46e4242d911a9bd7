// api.hip -- host runtime behind the C ABI declared in include/gpumesh.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpumesh.h"
#include "gm_device.h"
#include "gm_topology.h"

namespace
{

thread_local std::string g_lastError;

int fail(int code, const std::string& msg)
{
    g_lastError = msg;
    return code;
}

#define GM_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if(e_ != hipSuccess)                                                           \
            return fail(GM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while(0)

template<typename X> hipError_t devAlloc(X** p, size_t count)
{
    *p = nullptr;
    if(count == 0)
        count = 1;
    return hipMalloc((void**)p, count * sizeof(X));
}

} // namespace

struct gm_mesh
{
    bool onDevice = false;   // false: topology only (gm_mesh_create_host)
    int device = 0;
    hipStream_t stream = nullptr;
    gmt::Topology topo;
    gmd::DevMesh dm{};
    uint32_t* groupNodes = nullptr;
    int32_t* evalCounts = nullptr;
    void* getmeTet = nullptr;
    void* getmeHex = nullptr;
    size_t getmeBytes = 0;
    gmd::EvalScratch es{};
    int resultCap = 0;
    double* elemQual = nullptr;
    // crew
    int evaluator = gmq::EV_MEAN_RATIO;
    int sampler = gmq::SA_UNIFORM;
    gmq::Crew crew{};
    float* texels = nullptr;
    int maxPatch = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> sweepEv;
    double lastTotalMs = 0.0, lastSweepMs = 0.0;
    // Per-launch timing: event pairs around every kernel launch of a run,
    // tagged by kernel class (KT_*), summed after the run.
    std::vector<hipEvent_t> pool;
    std::vector<int> poolTag;
    size_t poolUsed = 0;
    double kernMs[GM_KT_COUNT] = {0};
    int64_t kernLaunches[GM_KT_COUNT] = {0};
};

namespace
{

// Opens a timed launch of class `tag`; returns the pool slot (or -1).
int timedBegin(gm_mesh* m, int tag)
{
    if(m->poolUsed + 2 > m->pool.size())
    {
        size_t old = m->pool.size();
        m->pool.resize(old + 256, nullptr);
        m->poolTag.resize(old + 256, -1);
        for(size_t i = old; i < m->pool.size(); ++i)
            if(hipEventCreate(&m->pool[i]) != hipSuccess) { m->pool.resize(i); m->poolTag.resize(i); return -1; }
        if(m->poolUsed + 2 > m->pool.size()) return -1;
    }
    int slot = (int)m->poolUsed;
    m->poolUsed += 2;
    m->poolTag[slot] = tag;
    (void)hipEventRecord(m->pool[slot], m->stream);
    return slot;
}

void timedEnd(gm_mesh* m, int slot)
{
    if(slot >= 0) (void)hipEventRecord(m->pool[slot + 1], m->stream);
}

// After the stream has drained: sums the durations per class.
void collectTimings(gm_mesh* m)
{
    for(int k = 0; k < GM_KT_COUNT; ++k) { m->kernMs[k] = 0.0; m->kernLaunches[k] = 0; }
    for(size_t i = 0; i + 1 < m->poolUsed; i += 2)
    {
        float ms = 0.0f;
        if(hipEventElapsedTime(&ms, m->pool[i], m->pool[i + 1]) == hipSuccess && m->poolTag[i] >= 0)
        {
            m->kernMs[m->poolTag[i]] += ms;
            m->kernLaunches[m->poolTag[i]] += 1;
        }
    }
    m->poolUsed = 0;
}

void freeMesh(gm_mesh* m)
{
    if(!m) return;
    if(!m->onDevice) { delete m; return; }
    (void)hipSetDevice(m->device);
    if(m->stream) (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->dm.pos64); (void)hipFree(m->dm.pos32); (void)hipFree((void*)m->dm.tets); (void)hipFree((void*)m->dm.hexs);
    (void)hipFree((void*)m->dm.eoff); (void)hipFree((void*)m->dm.eref); (void)hipFree((void*)m->dm.noff); (void)hipFree((void*)m->dm.nbr);
    (void)hipFree((void*)m->dm.fixed); (void)hipFree(m->groupNodes); (void)hipFree(m->evalCounts);
    (void)hipFree(m->getmeTet); (void)hipFree(m->getmeHex);
    (void)hipFree(m->es.partials); (void)hipFree(m->es.flags); (void)hipFree(m->es.result); (void)hipFree(m->es.hist);
    (void)hipFree(m->es.histOut); (void)hipFree(m->elemQual); (void)hipFree(m->texels);
    for(hipEvent_t e : m->ev) if(e) (void)hipEventDestroy(e);
    for(hipEvent_t e : m->sweepEv) if(e) (void)hipEventDestroy(e);
    for(hipEvent_t e : m->pool) if(e) (void)hipEventDestroy(e);
    if(m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int ensureResults(gm_mesh* m, int records)
{
    if(records <= m->resultCap)
        return GM_OK;
    (void)hipFree(m->es.result);
    (void)hipFree(m->es.histOut);
    m->es.result = nullptr;
    m->es.histOut = nullptr;
    GM_HIP(devAlloc(&m->es.result, 3 * (size_t)records));
    GM_HIP(devAlloc(&m->es.histOut, 20 * (size_t)records));
    m->resultCap = records;
    return GM_OK;
}

int ensureGetme(gm_mesh* m, size_t scalarBytes)
{
    size_t need = (4 * (size_t)m->topo.ntets + 8 * (size_t)m->topo.nhexs) * 4 * scalarBytes;
    if(need <= m->getmeBytes)
        return GM_OK;
    (void)hipFree(m->getmeTet);
    (void)hipFree(m->getmeHex);
    m->getmeTet = m->getmeHex = nullptr;
    size_t tetBytes = 4 * (size_t)m->topo.ntets * 4 * 8;   // sized for FP64
    size_t hexBytes = 8 * (size_t)m->topo.nhexs * 4 * 8;
    GM_HIP(hipMalloc(&m->getmeTet, tetBytes ? tetBytes : 32));
    GM_HIP(hipMalloc(&m->getmeHex, hexBytes ? hexBytes : 32));
    m->getmeBytes = tetBytes + hexBytes;
    return GM_OK;
}

int evaluate(gm_mesh* m, bool fromF32, int record)
{
    int slot = timedBegin(m, GM_KT_EVALUATE);
    hipError_t e = fromF32
        ? gmd::launchEvaluate<float>(m->dm, m->crew, m->evaluator, m->sampler, m->es, record, m->stream)
        : gmd::launchEvaluate<double>(m->dm, m->crew, m->evaluator, m->sampler, m->es, record, m->stream);
    timedEnd(m, slot);
    if(e != hipSuccess)
        return fail(GM_ERR_HIP, std::string("evaluate: ") + hipGetErrorString(e));
    return GM_OK;
}

} // namespace

extern "C" {

const char* gm_last_error(void) { return g_lastError.c_str(); }

const char* gm_version(void) { return "gpumesh_amd 0.1 (gfx950)"; }

int gm_device_count(int* count)
{
    if(!count) return fail(GM_ERR_ARG, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return GM_OK;
}

int gm_mesh_create(int64_t nverts, const double* pos, const int32_t* fixed,
                   int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                   gm_mesh** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(nverts < 0 || ntets < 0 || nhexs < 0) return fail(GM_ERR_ARG, "negative size");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs)) return fail(GM_ERR_ARG, "null array");
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GM_ERR_HIP, "no HIP device visible");

    gm_mesh* m = new gm_mesh();
    std::string err;
    if(!gmt::build(nverts, pos, fixed, ntets, tets, nhexs, hexs, m->topo, err))
    {
        delete m;
        return fail(GM_ERR_ARG, err);
    }
    const gmt::Topology& t = m->topo;
    for(int64_t v = 0; v < nverts; ++v)
        if(!t.fixed[v] && t.eoff[v + 1] == t.eoff[v])
        {
            delete m;
            return fail(GM_ERR_ARG, "a free vertex belongs to no element");
        }
    for(int32_t p : t.groupMaxPatch) m->maxPatch = std::max(m->maxPatch, (int)p);

    auto cleanupFail = [&](int code) { freeMesh(m); return code; };
#define GM_TRY(call) do { hipError_t e_ = (call); if(e_ != hipSuccess) { fail(GM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); return cleanupFail(e_ == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP); } } while(0)

    m->onDevice = true;
    GM_TRY(hipGetDevice(&m->device));
    GM_TRY(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    GM_TRY(hipEventCreate(&m->ev[0]));
    GM_TRY(hipEventCreate(&m->ev[1]));

    const size_t V = (size_t)nverts;
    std::vector<double> p64(4 * V);
    std::vector<float> p32(4 * V);
    for(size_t v = 0; v < V; ++v)
    {
        for(int d = 0; d < 3; ++d) { p64[4 * v + d] = t.pos[3 * v + d]; p32[4 * v + d] = (float)t.pos[3 * v + d]; }
        p64[4 * v + 3] = 0.0; p32[4 * v + 3] = 0.0f;
    }
    double4* pos64; float4* pos32; uint4* tb; uint4* hb; uint32_t *eo, *er, *no, *nb; uint8_t* fx;
    GM_TRY(devAlloc(&pos64, V)); m->dm.pos64 = pos64;
    GM_TRY(devAlloc(&pos32, V)); m->dm.pos32 = pos32;
    GM_TRY(devAlloc(&tb, (size_t)ntets)); m->dm.tets = tb;
    GM_TRY(devAlloc(&hb, 2 * (size_t)nhexs)); m->dm.hexs = hb;
    GM_TRY(devAlloc(&eo, V + 1)); m->dm.eoff = eo;
    GM_TRY(devAlloc(&er, t.eref.size())); m->dm.eref = er;
    GM_TRY(devAlloc(&no, V + 1)); m->dm.noff = no;
    GM_TRY(devAlloc(&nb, t.nbr.size())); m->dm.nbr = nb;
    GM_TRY(devAlloc(&fx, V)); m->dm.fixed = fx;
    GM_TRY(devAlloc(&m->groupNodes, t.groupNodes.size()));
    GM_TRY(devAlloc(&m->evalCounts, V));
    GM_TRY(devAlloc(&m->es.partials, 3 * (size_t)gmd::EVAL_BLOCKS));
    GM_TRY(devAlloc(&m->es.flags, (size_t)gmd::EVAL_BLOCKS));
    GM_TRY(devAlloc(&m->es.hist, 20));

    hipStream_t s = m->stream;
    if(V)
    {
        GM_TRY(hipMemcpyAsync(pos64, p64.data(), V * sizeof(double4), hipMemcpyHostToDevice, s));
        GM_TRY(hipMemcpyAsync(pos32, p32.data(), V * sizeof(float4), hipMemcpyHostToDevice, s));
        GM_TRY(hipMemcpyAsync(eo, t.eoff.data(), (V + 1) * 4, hipMemcpyHostToDevice, s));
        GM_TRY(hipMemcpyAsync(no, t.noff.data(), (V + 1) * 4, hipMemcpyHostToDevice, s));
        GM_TRY(hipMemcpyAsync(fx, t.fixed.data(), V, hipMemcpyHostToDevice, s));
    }
    if(ntets) GM_TRY(hipMemcpyAsync(tb, t.tets.data(), (size_t)ntets * 16, hipMemcpyHostToDevice, s));
    if(nhexs) GM_TRY(hipMemcpyAsync(hb, t.hexs.data(), (size_t)nhexs * 32, hipMemcpyHostToDevice, s));
    if(!t.eref.empty()) GM_TRY(hipMemcpyAsync(er, t.eref.data(), t.eref.size() * 4, hipMemcpyHostToDevice, s));
    if(!t.nbr.empty()) GM_TRY(hipMemcpyAsync(nb, t.nbr.data(), t.nbr.size() * 4, hipMemcpyHostToDevice, s));
    if(!t.groupNodes.empty())
        GM_TRY(hipMemcpyAsync(m->groupNodes, t.groupNodes.data(), t.groupNodes.size() * 4, hipMemcpyHostToDevice, s));
    GM_TRY(hipMemsetAsync(m->es.hist, 0, 20 * sizeof(int32_t), s));
    GM_TRY(hipStreamSynchronize(s));
#undef GM_TRY

    m->dm.nverts = (uint32_t)nverts;
    m->dm.ntets = (uint32_t)ntets;
    m->dm.nhexs = (uint32_t)nhexs;
    m->crew.scaling = m->crew.scaling2 = m->crew.scaling3 = 1.0;
    *out = m;
    return GM_OK;
}

int gm_mesh_create_host(int64_t nverts, const double* pos, const int32_t* fixed,
                        int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                        gm_mesh** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(nverts < 0 || ntets < 0 || nhexs < 0) return fail(GM_ERR_ARG, "negative size");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs)) return fail(GM_ERR_ARG, "null array");
    gm_mesh* m = new gm_mesh();
    std::string err;
    if(!gmt::build(nverts, pos, fixed, ntets, tets, nhexs, hexs, m->topo, err))
    {
        delete m;
        return fail(GM_ERR_ARG, err);
    }
    for(int32_t p : m->topo.groupMaxPatch) m->maxPatch = std::max(m->maxPatch, (int)p);
    *out = m;
    return GM_OK;
}

int gm_mesh_destroy(gm_mesh* mesh)
{
    freeMesh(mesh);
    return GM_OK;
}

int gm_mesh_info(const gm_mesh* m, int64_t* info)
{
    if(!m || !info) return fail(GM_ERR_ARG, "null argument");
    const gmt::Topology& t = m->topo;
    info[0] = t.nverts; info[1] = t.ntets; info[2] = t.nhexs;
    info[3] = (int64_t)t.groupNodes.size(); info[4] = t.ngroups();
    info[5] = (int64_t)t.eref.size(); info[6] = (int64_t)t.nbr.size(); info[7] = m->maxPatch;
    return GM_OK;
}

int gm_mesh_permutation(const gm_mesh* m, int64_t* perm)
{
    if(!m || !perm) return fail(GM_ERR_ARG, "null argument");
    std::memcpy(perm, m->topo.permutation.data(), m->topo.permutation.size() * sizeof(int64_t));
    return GM_OK;
}

int gm_mesh_groups(const gm_mesh* m, int64_t* off, uint32_t* nodes)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(off) std::memcpy(off, m->topo.groupOff.data(), m->topo.groupOff.size() * sizeof(int64_t));
    if(nodes) std::memcpy(nodes, m->topo.groupNodes.data(), m->topo.groupNodes.size() * 4);
    return GM_OK;
}

int gm_mesh_elements(const gm_mesh* m, uint32_t* tets, uint32_t* hexs)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(tets) std::memcpy(tets, m->topo.tets.data(), m->topo.tets.size() * 4);
    if(hexs) std::memcpy(hexs, m->topo.hexs.data(), m->topo.hexs.size() * 4);
    return GM_OK;
}

int gm_mesh_topology(const gm_mesh* m, uint32_t* eoff, uint32_t* eref, uint32_t* noff, uint32_t* nbr)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    const gmt::Topology& t = m->topo;
    if(eoff) std::memcpy(eoff, t.eoff.data(), t.eoff.size() * 4);
    if(eref) std::memcpy(eref, t.eref.data(), t.eref.size() * 4);
    if(noff) std::memcpy(noff, t.noff.data(), t.noff.size() * 4);
    if(nbr) std::memcpy(nbr, t.nbr.data(), t.nbr.size() * 4);
    return GM_OK;
}

int gm_mesh_get_positions(gm_mesh* m, double* pos)
{
    if(!m || !pos) return fail(GM_ERR_ARG, "null argument");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    GM_HIP(hipSetDevice(m->device));
    size_t V = (size_t)m->topo.nverts;
    std::vector<double> p64(4 * V);
    if(V) GM_HIP(hipMemcpyAsync(p64.data(), m->dm.pos64, V * sizeof(double4), hipMemcpyDeviceToHost, m->stream));
    GM_HIP(hipStreamSynchronize(m->stream));
    for(size_t v = 0; v < V; ++v) for(int d = 0; d < 3; ++d) pos[3 * v + d] = p64[4 * v + d];
    return GM_OK;
}

int gm_mesh_set_positions(gm_mesh* m, const double* pos)
{
    if(!m || !pos) return fail(GM_ERR_ARG, "null argument");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    GM_HIP(hipSetDevice(m->device));
    size_t V = (size_t)m->topo.nverts;
    std::vector<double> p64(4 * V);
    std::vector<float> p32(4 * V);
    for(size_t v = 0; v < V; ++v)
    {
        for(int d = 0; d < 3; ++d) { p64[4 * v + d] = pos[3 * v + d]; p32[4 * v + d] = (float)pos[3 * v + d]; }
        p64[4 * v + 3] = 0.0; p32[4 * v + 3] = 0.0f;
    }
    if(V)
    {
        GM_HIP(hipMemcpyAsync(m->dm.pos64, p64.data(), V * sizeof(double4), hipMemcpyHostToDevice, m->stream));
        GM_HIP(hipMemcpyAsync(m->dm.pos32, p32.data(), V * sizeof(float4), hipMemcpyHostToDevice, m->stream));
    }
    GM_HIP(hipStreamSynchronize(m->stream));
    return GM_OK;
}

int gm_mesh_set_crew(gm_mesh* m, int evaluator, int sampler, double scaling,
                     const int32_t* tsize, const double* tmin, const double* text, const float* texels)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(evaluator < 0 || evaluator > 1) return fail(GM_ERR_ARG, "unknown evaluator");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    if(sampler < 0 || sampler > 2) return fail(GM_ERR_ARG, "unknown sampler");
    if(!(scaling > 0.0)) return fail(GM_ERR_ARG, "scaling must be positive");
    GM_HIP(hipSetDevice(m->device));
    gmq::Crew c{};
    c.scaling = scaling;
    c.scaling2 = scaling * scaling;
    c.scaling3 = c.scaling2 * scaling;
    if(sampler == gmq::SA_TEXTURE)
    {
        if(!tsize || !tmin || !text || !texels) return fail(GM_ERR_ARG, "texture sampler needs a grid");
        for(int d = 0; d < 3; ++d)
        {
            if(tsize[d] < 1) return fail(GM_ERR_ARG, "texture size must be >= 1");
            if(!(text[d] > 0.0)) return fail(GM_ERR_ARG, "texture extents must be positive");
            c.tsize[d] = tsize[d]; c.tmin[d] = tmin[d]; c.text[d] = text[d];
        }
        size_t n = 6 * (size_t)tsize[0] * tsize[1] * tsize[2];
        (void)hipFree(m->texels);
        m->texels = nullptr;
        GM_HIP(devAlloc(&m->texels, n));
        GM_HIP(hipMemcpyAsync(m->texels, texels, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
        GM_HIP(hipStreamSynchronize(m->stream));
        c.texels = m->texels;
    }
    m->crew = c;
    m->evaluator = evaluator;
    m->sampler = sampler;
    return GM_OK;
}

int gm_mesh_evaluate(gm_mesh* m, double* out3, int32_t* hist20, double* elemQual)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    GM_HIP(hipSetDevice(m->device));
    int rc = ensureResults(m, 1);
    if(rc) return rc;
    size_t E = (size_t)(m->topo.ntets + m->topo.nhexs);
    if(elemQual && !m->elemQual) GM_HIP(devAlloc(&m->elemQual, E));
    m->es.elemQual = elemQual ? m->elemQual : nullptr;
    m->poolUsed = 0;
    rc = evaluate(m, false, 0);
    m->es.elemQual = nullptr;
    if(rc) return rc;
    double r[3];
    int32_t h[20];
    GM_HIP(hipMemcpyAsync(r, m->es.result, sizeof(r), hipMemcpyDeviceToHost, m->stream));
    GM_HIP(hipMemcpyAsync(h, m->es.histOut, sizeof(h), hipMemcpyDeviceToHost, m->stream));
    if(elemQual && E) GM_HIP(hipMemcpyAsync(elemQual, m->elemQual, E * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    GM_HIP(hipStreamSynchronize(m->stream));
    if(out3) std::memcpy(out3, r, sizeof(r));
    if(hist20) std::memcpy(hist20, h, sizeof(h));
    return GM_OK;
}

int gm_mesh_smooth(gm_mesh* m, int smoother, int passes, int precision, int kernel,
                   double spawnCoeff, double* stats, int32_t* hist, int32_t* evalCounts)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    if(passes < 0) return fail(GM_ERR_ARG, "negative pass count");
    if(precision < 0 || precision > 2) return fail(GM_ERR_ARG, "unknown precision");
    if(kernel < 0 || kernel > 1) return fail(GM_ERR_ARG, "unknown kernel shape");
    if(!(spawnCoeff > 0.0)) return fail(GM_ERR_ARG, "spawn coefficient must be positive");
    if(m->sampler == GM_TEXTURE && (kernel == GM_KERNEL_PLAIN || smoother == gmd::SM_GETME))
        return fail(GM_ERR_UNSUPPORTED, "the texture sampler runs only in the wavefront kernels of the vertex-wise "
                                        "smoothers (GM_KERNEL_WAVE with Nelder-Mead, gradient descent or spawn search)");
    GM_HIP(hipSetDevice(m->device));
    int rc = ensureResults(m, passes + 1);
    if(rc) return rc;
    if(smoother == gmd::SM_GETME)
    {
        rc = ensureGetme(m, 8);
        if(rc) return rc;
    }
    const gmt::Topology& t = m->topo;
    hipStream_t s = m->stream;
    size_t V = (size_t)t.nverts;
    if(V) GM_HIP(hipMemsetAsync(m->evalCounts, 0, V * sizeof(int32_t), s));
    if((int)m->sweepEv.size() < 2 * passes)
    {
        size_t old = m->sweepEv.size();
        m->sweepEv.resize(2 * passes, nullptr);
        for(size_t i = old; i < m->sweepEv.size(); ++i) GM_HIP(hipEventCreate(&m->sweepEv[i]));
    }
    m->poolUsed = 0;
    GM_HIP(hipEventRecord(m->ev[0], s));

    bool inF32 = false;
    if(precision != GM_FP64)
    {
        GM_HIP(gmd::launchToF32(m->dm, s));
        inF32 = true;
    }
    for(int pass = 0; pass <= passes; ++pass)
    {
        rc = evaluate(m, inF32, pass);
        if(rc) return rc;
        if(pass == passes)
            break;
        bool f64 = precision == GM_FP64 || (precision == GM_FP32_POLISH && pass == passes - 1);
        if(f64 && inF32)
        {
            GM_HIP(gmd::launchToF64(m->dm, s));
            inF32 = false;
        }
        GM_HIP(hipEventRecord(m->sweepEv[2 * pass], s));
        hipError_t e = hipSuccess;
        if(smoother == gmd::SM_GETME)
        {
            void* to = m->getmeTet;
            void* ho = m->getmeHex;
            int slot = timedBegin(m, f64 ? GM_KT_GETME_ELEM_F64 : GM_KT_GETME_ELEM_F32);
            e = f64 ? gmd::launchGetmeElements<double>(m->dm, m->crew, m->evaluator, m->sampler, to, ho, s)
                    : gmd::launchGetmeElements<float>(m->dm, m->crew, m->evaluator, m->sampler, to, ho, s);
            timedEnd(m, slot);
            for(int64_t g = 0; e == hipSuccess && g < t.ngroups(); ++g)
            {
                const uint32_t* nodes = m->groupNodes + t.groupOff[g];
                uint32_t cnt = (uint32_t)(t.groupOff[g + 1] - t.groupOff[g]);
                if(cnt == 0) continue;
                slot = timedBegin(m, f64 ? GM_KT_GETME_VERT_F64 : GM_KT_GETME_VERT_F32);
                e = f64 ? gmd::launchGetmeVertices<double>(m->dm, m->crew, m->evaluator, m->sampler, nodes, cnt, to, ho, m->evalCounts, s)
                        : gmd::launchGetmeVertices<float>(m->dm, m->crew, m->evaluator, m->sampler, nodes, cnt, to, ho, m->evalCounts, s);
                timedEnd(m, slot);
            }
        }
        else
        {
            for(int64_t g = 0; e == hipSuccess && g < t.ngroups(); ++g)
            {
                const uint32_t* nodes = m->groupNodes + t.groupOff[g];
                uint32_t cnt = (uint32_t)(t.groupOff[g + 1] - t.groupOff[g]);
                int mp = t.groupMaxPatch[g];
                if(cnt == 0) continue;
                int slot = timedBegin(m, f64 ? GM_KT_SMOOTH_F64 : GM_KT_SMOOTH_F32);
                e = f64 ? gmd::launchSmoothGroup<double>(m->dm, m->crew, m->evaluator, m->sampler, smoother, kernel, nodes, cnt, spawnCoeff, m->evalCounts, mp, s)
                        : gmd::launchSmoothGroup<float>(m->dm, m->crew, m->evaluator, m->sampler, smoother, kernel, nodes, cnt, spawnCoeff, m->evalCounts, mp, s);
                timedEnd(m, slot);
            }
        }
        if(e != hipSuccess)
            return fail(GM_ERR_HIP, std::string("smoothing launch: ") + hipGetErrorString(e));
        GM_HIP(hipEventRecord(m->sweepEv[2 * pass + 1], s));
    }
    if(inF32)
        GM_HIP(gmd::launchToF64(m->dm, s));
    GM_HIP(hipEventRecord(m->ev[1], s));

    std::vector<double> r(3 * (size_t)(passes + 1));
    std::vector<int32_t> h(20 * (size_t)(passes + 1));
    GM_HIP(hipMemcpyAsync(r.data(), m->es.result, r.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    GM_HIP(hipMemcpyAsync(h.data(), m->es.histOut, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if(evalCounts && V) GM_HIP(hipMemcpyAsync(evalCounts, m->evalCounts, V * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GM_HIP(hipStreamSynchronize(s));

    collectTimings(m);
    float ms = 0.0f;
    GM_HIP(hipEventElapsedTime(&ms, m->ev[0], m->ev[1]));
    m->lastTotalMs = ms;
    m->lastSweepMs = 0.0;
    for(int p = 0; p < passes; ++p)
    {
        GM_HIP(hipEventElapsedTime(&ms, m->sweepEv[2 * p], m->sweepEv[2 * p + 1]));
        m->lastSweepMs += ms;
    }
    for(int p = 0; p <= passes; ++p)
    {
        if(stats)
        {
            stats[3 * p] = r[3 * p];
            stats[3 * p + 1] = r[3 * p + 2] / r[3 * p + 1];
            stats[3 * p + 2] = r[3 * p + 1];
        }
        if(hist) std::memcpy(hist + 20 * p, h.data() + 20 * p, 20 * sizeof(int32_t));
    }
    return GM_OK;
}

int gm_mesh_smooth_groups(gm_mesh* m, int smoother, int64_t groupBegin, int64_t groupEnd, int precision,
                          int kernel, double spawnCoeff, int32_t* evalCounts)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(!m->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)");
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(smoother == gmd::SM_GETME)
        return fail(GM_ERR_UNSUPPORTED, "GETMe sweeps start with an element step over the whole mesh; use gm_mesh_smooth");
    if(precision != GM_FP64 && precision != GM_FP32) return fail(GM_ERR_ARG, "precision must be GM_FP64 or GM_FP32");
    if(kernel < 0 || kernel > 1) return fail(GM_ERR_ARG, "unknown kernel shape");
    if(!(spawnCoeff > 0.0)) return fail(GM_ERR_ARG, "spawn coefficient must be positive");
    if(kernel == GM_KERNEL_PLAIN && m->sampler == GM_TEXTURE)
        return fail(GM_ERR_UNSUPPORTED, "the texture sampler runs only in the wavefront kernels of the vertex-wise "
                                        "smoothers (GM_KERNEL_WAVE with Nelder-Mead, gradient descent or spawn search)");
    const gmt::Topology& t = m->topo;
    if(groupBegin < 0 || groupEnd > t.ngroups() || groupBegin > groupEnd) return fail(GM_ERR_ARG, "bad group range");
    GM_HIP(hipSetDevice(m->device));
    hipStream_t s = m->stream;
    size_t V = (size_t)t.nverts;
    if(V) GM_HIP(hipMemsetAsync(m->evalCounts, 0, V * sizeof(int32_t), s));
    const bool f64 = precision == GM_FP64;
    if(!f64) GM_HIP(gmd::launchToF32(m->dm, s));
    for(int64_t g = groupBegin; g < groupEnd; ++g)
    {
        const uint32_t* nodes = m->groupNodes + t.groupOff[g];
        uint32_t cnt = (uint32_t)(t.groupOff[g + 1] - t.groupOff[g]);
        hipError_t e = f64 ? gmd::launchSmoothGroup<double>(m->dm, m->crew, m->evaluator, m->sampler, smoother, kernel, nodes, cnt, spawnCoeff, m->evalCounts, t.groupMaxPatch[g], s)
                           : gmd::launchSmoothGroup<float>(m->dm, m->crew, m->evaluator, m->sampler, smoother, kernel, nodes, cnt, spawnCoeff, m->evalCounts, t.groupMaxPatch[g], s);
        if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("smoothing launch: ") + hipGetErrorString(e));
    }
    if(!f64) GM_HIP(gmd::launchToF64(m->dm, s));
    if(evalCounts && V) GM_HIP(hipMemcpyAsync(evalCounts, m->evalCounts, V * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GM_HIP(hipStreamSynchronize(s));
    return GM_OK;
}

int gm_mesh_last_timing(const gm_mesh* m, double* out2)
{
    if(!m || !out2) return fail(GM_ERR_ARG, "null argument");
    out2[0] = m->lastTotalMs;
    out2[1] = m->lastSweepMs;
    return GM_OK;
}

int gm_mesh_kernel_times(const gm_mesh* m, double* ms, int64_t* launches)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    for(int k = 0; k < GM_KT_COUNT; ++k)
    {
        if(ms) ms[k] = m->kernMs[k];
        if(launches) launches[k] = m->kernLaunches[k];
    }
    return GM_OK;
}

int gm_math_probe(int op, const double* a, const double* b, double* out, int64_t n)
{
    if(op < 0 || op > 4 || n < 0 || (n && (!a || !out || (op == 1 && !b)))) return fail(GM_ERR_ARG, "bad probe arguments");
    if(n == 0) return GM_OK;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GM_ERR_HIP, "no HIP device visible");
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    GM_HIP(devAlloc(&da, (size_t)n));
    GM_HIP(devAlloc(&db, (size_t)n));
    GM_HIP(devAlloc(&dout, (size_t)n));
    GM_HIP(hipMemcpy(da, a, n * sizeof(double), hipMemcpyHostToDevice));
    GM_HIP(hipMemcpy(db, b ? b : a, n * sizeof(double), hipMemcpyHostToDevice));
    GM_HIP(gmd::launchMathProbe(op, da, db, dout, (int)n, nullptr));
    GM_HIP(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(da); (void)hipFree(db); (void)hipFree(dout);
    return GM_OK;
}

} // extern "C"
