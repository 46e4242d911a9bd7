// api.hip -- the C ABI of include/gpumesh.h: mesh handles (a plan of one
// mesh each), batches (a plan of many) and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "gm_batr.h"
#include "gm_runtime.h"

using gmr::fail;

#define GM_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if(e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP,         \
                        std::string(#call) + ": " + hipGetErrorString(e_));            \
    } while(0)

#define GM_NEED_DEVICE(m)                                                                                   \
    do {                                                                                                    \
        if(!(m)) return fail(GM_ERR_ARG, "null mesh");                                                      \
        if(!(m)->onDevice) return fail(GM_ERR_ARG, "mesh has no device copy (created with gm_mesh_create_host)"); \
    } while(0)

// Meshes held by live batches (by address, never dereferenced here): a
// batch's descriptor table points into its members' device buffers, so a
// member must not be rebuilt under it.
static std::mutex g_memberMu;
static std::multiset<const gm_mesh*> g_batchMembers;

static bool inLiveBatch(const gm_mesh* m)
{
    std::lock_guard<std::mutex> lock(g_memberMu);
    return g_batchMembers.count(m) != 0;
}

struct gm_batch
{
    int device = 0;
    hipStream_t stream = nullptr;
    gmr::Plan* plan = nullptr;
    // Launch shape per smoother; unset (ppg 0) = the default shape.
    gmd::LaneShape lane[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
};

namespace
{

template<typename X> hipError_t devAlloc(X** p, size_t count)
{
    return gmr::poolAlloc((void**)p, (count ? count : 1) * sizeof(X));
}

int checkRun(int smoother, int passes, int precision, int kernel, double spawnCoeff)
{
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(passes < 0) return fail(GM_ERR_ARG, "negative pass count");
    if(precision < 0 || precision > 3) return fail(GM_ERR_ARG, "unknown precision");
    if(kernel < 0 || kernel > 1) return fail(GM_ERR_ARG, "unknown kernel shape");
    if(!(spawnCoeff > 0.0)) return fail(GM_ERR_ARG, "spawn coefficient must be positive");
    return GM_OK;
}

// Batch members share one crew: the same evaluator, sampler, scaling and
// aspect ratio, and under the texture sampler the same grid and texels.
bool sameCrew(const gm_mesh* a, const gm_mesh* b)
{
    if(a->evaluator != b->evaluator || a->sampler != b->sampler || a->crew.scaling != b->crew.scaling ||
       a->crew.aspect != b->crew.aspect)
        return false;
    if(a->sampler != GM_TEXTURE) return true;
    for(int k = 0; k < 3; ++k)
        if(a->crew.tsize[k] != b->crew.tsize[k] || a->crew.tmin[k] != b->crew.tmin[k] || a->crew.text[k] != b->crew.text[k])
            return false;
    return a->texHash == b->texHash;
}

// Default launch shape: one wavefront per workgroup (GETMe vertex update: 4)
// and the LDS tile chosen per work list from its largest patch (tile 0), so
// that a list picks the same kernel alone and batched.
gmd::LaneShape defaultLane(const std::vector<gm_mesh*>&, int smoother)
{
    return gmd::LaneShape{smoother == gmd::SM_GETME ? 4 : 1, 0};
}

// Runs the plan and copies each member's records (stats/hist strided by
// passes + 1 records per member); passesRun[i] = sweeps member i ran.
int runAndFetch(gmr::Plan* p, const gm_mesh* m0, const gmr::RunParams& r, double* stats, int32_t* hist,
                int32_t* evalCounts, int32_t* passesRun = nullptr)
{
    int rc = gmr::runPlan(p, m0->crew, m0->evaluator, m0->sampler, r);
    if(rc) return rc;
    const size_t per = (size_t)r.passes + 1;
    for(size_t i = 0; i < p->meshes.size(); ++i)
    {
        int records = (int)per;
        if(r.autopilot && (rc = gmr::recordsWritten(p, i, &records))) return rc;
        if((rc = gmr::fetchRecords(p, i, records, stats ? stats + i * 3 * per : nullptr,
                                   hist ? hist + i * 20 * per : nullptr, evalCounts)))
            return rc;
        if(passesRun) passesRun[i] = records - 1;
    }
    return GM_OK;
}

int scheduleParams(const gm_schedule* sched, int precision, gmr::RunParams& r)
{
    if(!sched) return fail(GM_ERR_ARG, "schedule is null");
    if(!sched->autopilot) return GM_OK;
    if(precision == GM_FP32_POLISH)
        return fail(GM_ERR_UNSUPPORTED, "the autopilot decides after each pass whether another runs, so no pass "
                                        "is known to be the last one (use GM_FP64 or GM_FP32)");
    if(!(sched->min_qual_threshold == sched->min_qual_threshold) ||
       !(sched->qual_mean_threshold == sched->qual_mean_threshold))
        return fail(GM_ERR_ARG, "autopilot thresholds must not be NaN");
    r.autopilot = true;
    r.apMeanGain = sched->min_qual_threshold;
    r.apMinGain = sched->qual_mean_threshold;
    return GM_OK;
}

double sweepMs(const gmr::Plan* p)
{
    double s = 0.0;
    for(int k = 0; k < GM_KT_EVALUATE; ++k) s += p->kernMs[k];
    return s;
}

} // namespace

extern "C" {

const char* gm_last_error(void) { return gmr::lastError(); }

const char* gm_version(void) { return "gpumesh_amd 0.2 (gfx950)"; }

int gm_device_count(int* count)
{
    if(!count) return fail(GM_ERR_ARG, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return GM_OK;
}

int gm_mesh_create_host(int64_t nverts, const double* pos, const int32_t* fixed,
                        int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                        gm_mesh** out)
{
    return gmr::makeMesh(nverts, pos, fixed, ntets, tets, nhexs, hexs, out);
}

int gm_mesh_create(int64_t nverts, const double* pos, const int32_t* fixed,
                   int64_t ntets, const uint32_t* tets, int64_t nhexs, const uint32_t* hexs,
                   gm_mesh** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GM_ERR_HIP, "no HIP device visible");
    gm_mesh* m = nullptr;
    int rc = gmr::makeMesh(nverts, pos, fixed, ntets, tets, nhexs, hexs, &m);
    if(rc) return rc;
    if((rc = gmr::uploadMesh(m)))
    {
        gmr::destroyMesh(m);
        return rc;
    }
    *out = m;
    return GM_OK;
}

int gm_mesh_create_grouped(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                           const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int64_t ngroups,
                           const int64_t* group_off, const uint32_t* group_nodes, int device_copy, gm_mesh** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(device_copy)
    {
        int ndev = 0;
        if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GM_ERR_HIP, "no HIP device visible");
    }
    gm_mesh* m = nullptr;
    int rc = gmr::makeMeshGrouped(nverts, pos, fixed, ntets, tets, nhexs, hexs, ngroups, group_off, group_nodes, &m);
    if(rc) return rc;
    if(device_copy && (rc = gmr::uploadMesh(m)))
    {
        gmr::destroyMesh(m);
        return rc;
    }
    *out = m;
    return GM_OK;
}

int gm_mesh_create_desc(const gm_mesh_desc* d, gm_mesh** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(!d) return fail(GM_ERR_ARG, "null description");
    if(d->nconstraints < 0 || (d->nconstraints > 0 && !d->constraints))
        return fail(GM_ERR_ARG, "bad constraint table");
    if(d->npris < 0 || (d->npris > 0 && !d->pris)) return fail(GM_ERR_ARG, "bad prism array");
    // The table in the device layout: [kind, a, unit normal or direction, b,
    // length, 0]; the normalisations are the constraints' constructors'
    // (FaceConstraint.cpp:173-178, EdgeConstraint.cpp:177-183).
    std::vector<double> table((size_t)d->nconstraints * gmd::CONS_STRIDE, 0.0);
    for(int64_t k = 0; k < d->nconstraints; ++k)
    {
        const gm_constraint& c = d->constraints[k];
        double* e = &table[(size_t)k * gmd::CONS_STRIDE];
        double n[3];
        if(c.kind == GM_CONSTRAINT_PLANE)
            for(int i = 0; i < 3; ++i) n[i] = c.b[i];
        else if(c.kind == GM_CONSTRAINT_SEGMENT)
            for(int i = 0; i < 3; ++i) n[i] = c.b[i] - c.a[i];
        else
            return fail(GM_ERR_ARG, "constraint " + std::to_string(k) + " has an unknown kind");
        const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if(!(len > 0.0) || len == (double)__builtin_huge_val())
            return fail(GM_ERR_ARG, "constraint " + std::to_string(k) + " has a zero or non-finite normal or length");
        const double inv = 1.0 / len;
        e[0] = (double)c.kind;
        for(int i = 0; i < 3; ++i)
        {
            e[1 + i] = c.a[i];
            e[4 + i] = n[i] * inv;
            e[7 + i] = c.b[i];
        }
        e[10] = len;
    }
    gm_mesh* m = nullptr;
    const std::vector<double>* tp = d->nconstraints > 0 ? &table : nullptr;
    int rc = d->ngroups < 0
                 ? gmr::makeMesh(d->nverts, d->pos, d->fixed, d->ntets, d->tets, d->nhexs, d->hexs, &m,
                                 d->constraint, tp, d->npris, d->pris)
                 : gmr::makeMeshGrouped(d->nverts, d->pos, d->fixed, d->ntets, d->tets, d->nhexs, d->hexs, d->ngroups,
                                        d->group_off, d->group_nodes, &m, d->constraint, tp, d->npris, d->pris);
    if(rc) return rc;
    if(d->device_copy)
    {
        int ndev = 0;
        if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        {
            gmr::destroyMesh(m);
            return fail(GM_ERR_HIP, "no HIP device visible");
        }
        if((rc = gmr::uploadMesh(m)))
        {
            gmr::destroyMesh(m);
            return rc;
        }
    }
    *out = m;
    return GM_OK;
}

int gm_mesh_destroy(gm_mesh* mesh)
{
    gmr::destroyMesh(mesh);
    return GM_OK;
}

int gm_mesh_info(const gm_mesh* m, int64_t* info)
{
    if(!m || !info) return fail(GM_ERR_ARG, "null argument");
    const gmt::Topology& t = *m->topo;
    info[0] = t.nverts;
    info[1] = t.ntets;
    info[2] = t.nhexs;
    info[3] = (int64_t)t.groupNodes.size();
    info[4] = t.ngroups();
    info[5] = (int64_t)t.eref.size();
    info[6] = (int64_t)t.nbr.size();
    info[7] = m->maxPatch;
    return GM_OK;
}

int gm_set_topology_cache(int64_t max_bytes)
{
    if(max_bytes < 0) return fail(GM_ERR_ARG, "negative cache size");
    gmt::setCacheCapacity((size_t)max_bytes);
    return GM_OK;
}

int gm_set_polish_passes(int passes)
{
    if(passes < 1 || passes > 64) return fail(GM_ERR_ARG, "polish passes must be in [1, 64]");
    gmr::setPolishPasses(passes);
    return GM_OK;
}

int gm_set_device_pool(int64_t max_cached_bytes)
{
    if(max_cached_bytes < 0) return fail(GM_ERR_ARG, "negative pool size");
    gmr::setPoolCap((size_t)max_cached_bytes);
    return GM_OK;
}

int gm_topology_prepare(int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                        const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int device_copy)
{
    if(gmt::cacheCapacity() == 0) return fail(GM_ERR_ARG, "the topology cache is off (gm_set_topology_cache)");
    gm_mesh* m = nullptr;
    int rc = gmr::makeMesh(nverts, pos, fixed, ntets, tets, nhexs, hexs, &m);
    if(rc) return rc;
    if(!gmt::isCached(m->topo.get()))
        rc = fail(GM_ERR_NOMEM, "the topology is larger than the topology cache's capacity");
    if(!rc && device_copy)
    {
        int ndev = 0;
        if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) rc = fail(GM_ERR_HIP, "no HIP device visible");
        else rc = gmr::prepareDeviceTopology(m);
    }
    gmr::destroyMesh(m);
    return rc;
}

int gm_topology_cache_stats(int64_t* out4)
{
    if(!out4) return fail(GM_ERR_ARG, "null argument");
    gmt::cacheStats(out4);
    return GM_OK;
}

int gm_mesh_device_bytes(const gm_mesh* m, int64_t* bytes)
{
    if(!m || !bytes) return fail(GM_ERR_ARG, "null argument");
    *bytes = m->onDevice ? (int64_t)gmr::residentBytes(m, m->resultCap, m->getmeReady) : 0;
    return GM_OK;
}

int gm_mesh_permutation(const gm_mesh* m, int64_t* perm)
{
    if(!m || !perm) return fail(GM_ERR_ARG, "null argument");
    std::memcpy(perm, m->topo->permutation.data(), m->topo->permutation.size() * sizeof(int64_t));
    return GM_OK;
}

int gm_mesh_groups(const gm_mesh* m, int64_t* off, uint32_t* nodes)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(off) std::memcpy(off, m->topo->groupOff.data(), m->topo->groupOff.size() * sizeof(int64_t));
    if(nodes) std::memcpy(nodes, m->topo->groupNodes.data(), m->topo->groupNodes.size() * 4);
    return GM_OK;
}

int gm_mesh_elements(const gm_mesh* m, uint32_t* tets, uint32_t* hexs)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(tets) std::memcpy(tets, m->topo->tets.data(), m->topo->tets.size() * 4);
    if(hexs) std::memcpy(hexs, m->topo->hexs.data(), 8 * (size_t)m->topo->nhexs * 4);
    return GM_OK;
}

int gm_mesh_prisms(const gm_mesh* m, int64_t* npris, uint32_t* pris)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    const gmt::Topology& t = *m->topo;
    if(npris) *npris = t.npris;
    if(pris)
        for(int64_t p = 0; p < t.npris; ++p)
            for(int k = 0; k < 6; ++k) pris[6 * p + k] = t.hexs[8 * (size_t)(t.nhexs + p) + k];
    return GM_OK;
}

int gm_mesh_topology(const gm_mesh* m, uint32_t* eoff, uint32_t* eref, uint32_t* noff, uint32_t* nbr)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    const gmt::Topology& t = *m->topo;
    if(eoff) std::memcpy(eoff, t.eoff.data(), t.eoff.size() * 4);
    if(eref) std::memcpy(eref, t.eref.data(), t.eref.size() * 4);
    if(noff) std::memcpy(noff, t.noff.data(), t.noff.size() * 4);
    if(nbr) std::memcpy(nbr, t.nbr.data(), t.nbr.size() * 4);
    return GM_OK;
}

int gm_mesh_get_positions(gm_mesh* m, double* pos)
{
    if(!m || !pos) return fail(GM_ERR_ARG, "null argument");
    if(!m->onDevice)
    {
        std::copy(m->hostPos.begin(), m->hostPos.end(), pos);
        return GM_OK;
    }
    GM_HIP(hipSetDevice(m->device));
    const size_t V = (size_t)m->topo->nverts;
    std::vector<double> p64(4 * V);
    if(V) GM_HIP(hipMemcpyAsync(p64.data(), m->dm.pos64, V * sizeof(double4), hipMemcpyDeviceToHost, m->stream));
    GM_HIP(hipStreamSynchronize(m->stream));
    for(size_t v = 0; v < V; ++v)
        for(int k = 0; k < 3; ++k) pos[3 * v + k] = p64[4 * (size_t)m->slotOf[v] + k];
    return GM_OK;
}

int gm_mesh_set_positions(gm_mesh* m, const double* pos)
{
    GM_NEED_DEVICE(m);
    if(!pos) return fail(GM_ERR_ARG, "null argument");
    GM_HIP(hipSetDevice(m->device));
    const size_t V = (size_t)m->topo->nverts;
    std::vector<double> p64(4 * V);
    std::vector<float> p32(4 * V);
    for(size_t v = 0; v < V; ++v)
    {
        const size_t s = m->slotOf[v];
        for(int k = 0; k < 3; ++k)
        {
            p64[4 * s + k] = pos[3 * v + k];
            p32[4 * s + k] = (float)pos[3 * v + k];
        }
        p64[4 * s + 3] = 0.0;
        p32[4 * s + 3] = 0.0f;
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
    if(sampler < 0 || sampler > 3) return fail(GM_ERR_ARG, "unknown sampler");
    if(!(scaling > 0.0)) return fail(GM_ERR_ARG, "scaling must be positive");
    if(!m->onDevice && sampler == GM_TEXTURE)
        return fail(GM_ERR_ARG, "the texture sampler needs a device copy (mesh created with gm_mesh_create_host)");
    if(m->onDevice) GM_HIP(hipSetDevice(m->device));
    gmq::Crew c{};
    c.scaling = scaling;
    c.scaling2 = scaling * scaling;
    c.scaling3 = c.scaling2 * scaling;
    gmr::setAnalyticField(c, m->aspect);
    if(sampler == GM_TEXTURE)
    {
        if(!tsize || !tmin || !text || !texels) return fail(GM_ERR_ARG, "texture sampler needs a grid");
        for(int k = 0; k < 3; ++k)
        {
            if(tsize[k] < 1) return fail(GM_ERR_ARG, "texture size must be >= 1");
            if(!(text[k] > 0.0)) return fail(GM_ERR_ARG, "texture extents must be positive");
            c.tsize[k] = tsize[k];
            c.tmin[k] = tmin[k];
            c.text[k] = text[k];
        }
        gmq::textureFields(c);
        // The new grid is allocated and filled before the old one is let go,
        // so a failure leaves the previous crew (and its texels) intact.
        const size_t n = 6 * (size_t)tsize[0] * tsize[1] * tsize[2];
        float* fresh = nullptr;
        GM_HIP(devAlloc(&fresh, n));
        hipError_t e = hipMemcpyAsync(fresh, texels, n * sizeof(float), hipMemcpyHostToDevice, m->stream);
        if(e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if(e != hipSuccess)
        {
            gmr::poolFree(fresh);
            return fail(GM_ERR_HIP, std::string("texture upload: ") + hipGetErrorString(e));
        }
        c.texels = fresh;
        gmr::poolFree(m->texels);
        m->texels = fresh;
        m->texHash = gmr::texelHash(texels, n);
    }
    m->crew = c;
    m->evaluator = evaluator;
    m->sampler = sampler;
    return GM_OK;
}

int gm_mesh_set_aspect_ratio(gm_mesh* m, double aspect)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(!(aspect > 0.0) || aspect == (double)__builtin_huge_val()) return fail(GM_ERR_ARG, "aspect ratio must be positive and finite");
    m->aspect = aspect;
    m->crew.aspect = aspect;
    return GM_OK;
}

int gm_mesh_evaluate(gm_mesh* m, double* out3, int32_t* hist20, double* elemQual)
{
    GM_NEED_DEVICE(m);
    GM_HIP(hipSetDevice(m->device));
    const size_t E = (size_t)(m->topo->ntets + m->topo->nhexs + m->topo->npris);
    if(elemQual && !m->elemQualBuf) GM_HIP(devAlloc(&m->elemQualBuf, E));
    gmr::Plan* p = m->plan;
    // Per-element output is switched on in the descriptor for this call only.
    m->dm.elemQual = elemQual ? m->elemQualBuf : nullptr;
    int rc = gmr::preparePlan(p, 1, false);
    m->dm.elemQual = nullptr;
    if(rc) return rc;
    GM_HIP(gmd::launchResetRun(p->dTable, p->vertBlocks, p->meshList, p->nMeshes, p->stream));
    hipError_t e = gmd::launchEvaluate<double>(p->dTable, p->evalBlocks, p->meshList, p->nMeshes, m->crew,
                                               m->evaluator, m->sampler, gmr::planHasRecords(p), p->stream);
    if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("evaluate: ") + hipGetErrorString(e));
    double r[3];
    int32_t h[20];
    GM_HIP(hipMemcpyAsync(r, m->dm.evalResult, sizeof(r), hipMemcpyDeviceToHost, p->stream));
    GM_HIP(hipMemcpyAsync(h, m->dm.evalHistOut, sizeof(h), hipMemcpyDeviceToHost, p->stream));
    if(elemQual && E)
        GM_HIP(hipMemcpyAsync(elemQual, m->elemQualBuf, E * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    GM_HIP(hipStreamSynchronize(p->stream));
    if(out3) std::memcpy(out3, r, sizeof(r));
    if(hist20) std::memcpy(hist20, h, sizeof(h));
    return GM_OK;
}

int gm_mesh_smooth(gm_mesh* m, int smoother, int passes, int precision, int kernel,
                   double spawnCoeff, double* stats, int32_t* hist, int32_t* evalCounts)
{
    GM_NEED_DEVICE(m);
    int rc = checkRun(smoother, passes, precision, kernel, spawnCoeff);
    if(rc) return rc;
    GM_HIP(hipSetDevice(m->device));
    gmr::RunParams r;
    r.smoother = smoother;
    r.passes = passes;
    r.precision = precision;
    r.polishPasses = gmr::polishPasses();
    r.kernel = kernel;
    r.spawnCoeff = spawnCoeff;
    r.lane = defaultLane(m->plan->meshes, smoother);
    return runAndFetch(m->plan, m, r, stats, hist, evalCounts);
}

int gm_mesh_smooth_schedule(gm_mesh* m, int smoother, int maxPasses, const gm_schedule* sched, int precision,
                            int kernel, double spawnCoeff, double* stats, int32_t* hist, int32_t* evalCounts,
                            int32_t* passesRun)
{
    GM_NEED_DEVICE(m);
    int rc = checkRun(smoother, maxPasses, precision, kernel, spawnCoeff);
    if(rc) return rc;
    GM_HIP(hipSetDevice(m->device));
    gmr::RunParams r;
    r.smoother = smoother;
    r.passes = maxPasses;
    r.precision = precision;
    r.polishPasses = gmr::polishPasses();
    r.kernel = kernel;
    r.spawnCoeff = spawnCoeff;
    r.lane = defaultLane(m->plan->meshes, smoother);
    if((rc = scheduleParams(sched, precision, r))) return rc;
    int32_t run = 0;
    rc = runAndFetch(m->plan, m, r, stats, hist, evalCounts, &run);
    if(!rc && passesRun) *passesRun = run;
    return rc;
}

int gm_mesh_restructure(gm_mesh* m, int topoPasses, int refinementSweeps, double minEdge, double maxEdge,
                        int64_t* counts)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(topoPasses < 1 || refinementSweeps < 0) return fail(GM_ERR_ARG, "topo_passes must be >= 1, refinement_sweeps >= 0");
    if(!(minEdge > 0.0) || !(maxEdge > minEdge) || maxEdge == (double)__builtin_huge_val())
        return fail(GM_ERR_ARG, "edge bounds must satisfy 0 < min_edge < max_edge < inf");
    if(counts)
        for(int k = 0; k < 5; ++k) counts[k] = 0;
    const gmt::Topology& t = *m->topo;
    if(t.ntets == 0 || t.nhexs || t.npris) return GM_OK;
    if(!t.cons.empty())
        return fail(GM_ERR_UNSUPPORTED, "topology operations need free or fixed vertices (this mesh has sliding ones)");
    if(inLiveBatch(m))
        return fail(GM_ERR_ARG, "the mesh is a member of a live batch: destroy the batch before restructuring");
    if(m->onDevice) GM_HIP(hipSetDevice(m->device));

    gmb::TetMesh in;
    in.pos.resize(3 * (size_t)t.nverts);
    int rc = gm_mesh_get_positions(m, in.pos.data());
    if(rc) return rc;
    // A non-finite coordinate makes every quality comparison false, and the
    // swaps (taken unless the new configuration is no better) would not end.
    for(double x : in.pos)
        if(!(x - x == 0.0)) return fail(GM_ERR_ARG, "restructure: a vertex position is not finite");
    in.fixed = t.fixed;
    in.tets = t.tets;
    in.eoff = t.eoff;
    in.eref = t.eref;
    in.noff = t.noff;
    in.nbr = t.nbr;
    gmb::Settings s;
    s.topoPasses = topoPasses;
    s.refinementSweeps = refinementSweeps;
    s.minEdge = minEdge;
    s.maxEdge = maxEdge;
    s.evaluator = m->evaluator;
    s.sampler = m->sampler;
    s.crew = m->crew;
    std::vector<float> texels;
    if(m->sampler == GM_TEXTURE && m->onDevice)
    {
        texels.resize(6 * (size_t)m->crew.tsize[0] * m->crew.tsize[1] * m->crew.tsize[2]);
        GM_HIP(hipMemcpyAsync(texels.data(), m->texels, texels.size() * sizeof(float), hipMemcpyDeviceToHost,
                              m->stream));
        GM_HIP(hipStreamSynchronize(m->stream));
        s.crew.texels = texels.data();
    }
    gmb::TetMesh out;
    gmb::Counts c;
    std::string err;
    if(!gmb::restructure(in, s, out, c, err)) return fail(GM_ERR_ARG, "restructure: " + err);

    // The new mesh: compiled (neighbourhoods, node groups, renumbering) and
    // uploaded beside the old one, which it replaces only once complete.
    const int64_t V = (int64_t)out.fixed.size();
    std::vector<int32_t> fixed(out.fixed.begin(), out.fixed.end());
    gm_mesh* fresh = nullptr;
    rc = gmr::makeMesh(V, out.pos.data(), fixed.data(), (int64_t)(out.tets.size() / 4), out.tets.data(), 0, nullptr,
                       &fresh);
    if(rc) return rc;
    fresh->evaluator = m->evaluator;
    fresh->sampler = m->sampler;
    fresh->aspect = m->aspect;
    fresh->crew = m->crew;
    fresh->texHash = m->texHash;
    if(m->onDevice)
    {
        rc = gmr::uploadMesh(fresh);
        if(rc)
        {
            gmr::destroyMesh(fresh);
            return rc;
        }
        // The texture grid moves over to the new mesh.
        fresh->texels = m->texels;
        m->texels = nullptr;
    }
    std::swap(*m, *fresh);
    if(m->plan) m->plan->meshes[0] = m;
    gmr::destroyMesh(fresh);
    if(counts)
    {
        counts[0] = c.faceSwaps;
        counts[1] = c.edgeSwaps;
        counts[2] = c.splits;
        counts[3] = c.merges;
        counts[4] = c.passes;
    }
    return GM_OK;
}

int gm_mesh_smooth_groups(gm_mesh* m, int smoother, int64_t groupBegin, int64_t groupEnd, int precision,
                          int kernel, double spawnCoeff, int32_t* evalCounts)
{
    GM_NEED_DEVICE(m);
    int rc = checkRun(smoother, 0, precision, kernel, spawnCoeff);
    if(rc) return rc;
    if(smoother == gmd::SM_GETME && !(groupBegin == 0 && groupEnd == m->topo->ngroups()))
        return fail(GM_ERR_UNSUPPORTED, "a GETMe sweep starts with an element step over the whole mesh: it runs "
                                        "only over all colour groups [0, ngroups)");
    if(precision == GM_FP32_POLISH)
        return fail(GM_ERR_ARG, "precision must be GM_FP64, GM_FP32 or GM_FP64_FAST (a partial sweep has no last pass)");
    if(groupBegin < 0 || groupEnd > m->topo->ngroups() || groupBegin > groupEnd)
        return fail(GM_ERR_ARG, "bad group range");
    GM_HIP(hipSetDevice(m->device));
    gmr::RunParams r;
    r.smoother = smoother;
    r.precision = precision;
    r.polishPasses = gmr::polishPasses();
    r.kernel = kernel;
    r.spawnCoeff = spawnCoeff;
    r.lane = defaultLane(m->plan->meshes, smoother);
    if((rc = gmr::sweepColours(m->plan, m->crew, m->evaluator, m->sampler, r, groupBegin, groupEnd))) return rc;
    if(evalCounts && m->topo->nverts)
    {
        const uint32_t V = (uint32_t)m->topo->nverts;
        GM_HIP(gmd::launchGatherCounts(m->dm.evalCounts, m->dSlotOf, m->dCountsRef, V, m->stream));
        GM_HIP(hipMemcpyAsync(evalCounts, m->dCountsRef, (size_t)V * 4, hipMemcpyDeviceToHost, m->stream));
        GM_HIP(hipStreamSynchronize(m->stream));
    }
    return GM_OK;
}

int gm_mesh_last_timing(const gm_mesh* m, double* out2)
{
    if(!m || !out2) return fail(GM_ERR_ARG, "null argument");
    if(!m->plan) return fail(GM_ERR_ARG, "mesh has no device copy");
    out2[0] = m->plan->totalMs;
    out2[1] = sweepMs(m->plan);
    return GM_OK;
}

int gm_mesh_kernel_times(const gm_mesh* m, double* ms, int64_t* launches)
{
    if(!m) return fail(GM_ERR_ARG, "null mesh");
    if(!m->plan) return fail(GM_ERR_ARG, "mesh has no device copy");
    for(int k = 0; k < GM_KT_COUNT; ++k)
    {
        if(ms) ms[k] = m->plan->kernMs[k];
        if(launches) launches[k] = m->plan->kernLaunches[k];
    }
    return GM_OK;
}

// ----- batches

int gm_batch_create(gm_mesh* const* meshes, int64_t n, gm_batch** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(n < 1 || !meshes) return fail(GM_ERR_ARG, "a batch needs at least one mesh");
    std::vector<gm_mesh*> ms(meshes, meshes + n);
    for(gm_mesh* m : ms)
    {
        if(!m || !m->onDevice) return fail(GM_ERR_ARG, "batch member without a device copy");
        if(m->device != ms[0]->device) return fail(GM_ERR_ARG, "batch members live on different devices");
        if(!sameCrew(m, ms[0]))
            return fail(GM_ERR_ARG, "batch members need the same evaluator, sampler, scaling and aspect ratio "
                                    "(and the same texture grid and texels under the texture sampler)");
    }
    std::vector<gm_mesh*> sorted = ms;
    std::sort(sorted.begin(), sorted.end());
    if(std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(GM_ERR_ARG, "a mesh appears twice in the batch");
    GM_HIP(hipSetDevice(ms[0]->device));
    gm_batch* b = new gm_batch();
    b->device = ms[0]->device;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if(e != hipSuccess)
    {
        delete b;
        return fail(GM_ERR_HIP, std::string("stream creation: ") + hipGetErrorString(e));
    }
    int rc = gmr::buildPlan(ms, b->stream, &b->plan);
    if(rc)
    {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lock(g_memberMu);
        for(const gm_mesh* m : ms) g_batchMembers.insert(m);
    }
    *out = b;
    return GM_OK;
}

int gm_batch_destroy(gm_batch* b)
{
    if(!b) return GM_OK;
    {
        std::lock_guard<std::mutex> lock(g_memberMu);
        for(const gm_mesh* m : b->plan->meshes)
        {
            auto it = g_batchMembers.find(m);
            if(it != g_batchMembers.end()) g_batchMembers.erase(it);
        }
    }
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    gmr::freePlan(b->plan);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return GM_OK;
}

int gm_batch_set_shape(gm_batch* b, int smoother, int patchesPerGroup, int tileElems)
{
    if(!b) return fail(GM_ERR_ARG, "null batch");
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(patchesPerGroup != 1 && patchesPerGroup != 2 && patchesPerGroup != 4 && patchesPerGroup != 8)
        return fail(GM_ERR_ARG, "patches per workgroup must be 1, 2, 4 or 8");
    if(tileElems != 64 && tileElems != 128 && tileElems != 256)
        return fail(GM_ERR_ARG, "LDS tile must be 64, 128 or 256 elements");
    b->lane[smoother] = gmd::LaneShape{patchesPerGroup, tileElems};
    return GM_OK;
}

static int checkBatchCrew(const gm_batch* b)
{
    const std::vector<gm_mesh*>& ms = b->plan->meshes;
    for(const gm_mesh* m : ms)
        if(!sameCrew(m, ms[0]))
            return fail(GM_ERR_ARG, "batch members' crews diverged after the batch was made");
    return GM_OK;
}

int gm_batch_smooth(gm_batch* b, int smoother, int passes, int precision, int kernel, double spawnCoeff,
                    int useGraph, double* stats, int32_t* hist)
{
    if(!b) return fail(GM_ERR_ARG, "null batch");
    int rc = checkRun(smoother, passes, precision, kernel, spawnCoeff);
    if(rc) return rc;
    if((rc = checkBatchCrew(b))) return rc;
    GM_HIP(hipSetDevice(b->device));
    gmr::RunParams r;
    r.smoother = smoother;
    r.passes = passes;
    r.precision = precision;
    r.polishPasses = gmr::polishPasses();
    r.kernel = kernel;
    r.spawnCoeff = spawnCoeff;
    r.lane = b->lane[smoother].patchesPerGroup ? b->lane[smoother] : defaultLane(b->plan->meshes, smoother);
    r.graph = useGraph != 0;
    return runAndFetch(b->plan, b->plan->meshes[0], r, stats, hist, nullptr);
}

int gm_batch_smooth_schedule(gm_batch* b, int smoother, int maxPasses, const gm_schedule* sched, int precision,
                             int kernel, double spawnCoeff, int useGraph, double* stats, int32_t* hist,
                             int32_t* passesRun)
{
    if(!b) return fail(GM_ERR_ARG, "null batch");
    int rc = checkRun(smoother, maxPasses, precision, kernel, spawnCoeff);
    if(rc) return rc;
    if((rc = checkBatchCrew(b))) return rc;
    GM_HIP(hipSetDevice(b->device));
    gmr::RunParams r;
    r.smoother = smoother;
    r.passes = maxPasses;
    r.precision = precision;
    r.polishPasses = gmr::polishPasses();
    r.kernel = kernel;
    r.spawnCoeff = spawnCoeff;
    r.lane = b->lane[smoother].patchesPerGroup ? b->lane[smoother] : defaultLane(b->plan->meshes, smoother);
    r.graph = useGraph != 0;
    if((rc = scheduleParams(sched, precision, r))) return rc;
    std::vector<int32_t> run(b->plan->meshes.size(), 0);
    rc = runAndFetch(b->plan, b->plan->meshes[0], r, stats, hist, nullptr, run.data());
    if(!rc && passesRun) std::memcpy(passesRun, run.data(), run.size() * sizeof(int32_t));
    return rc;
}

int gm_batch_last_timing(const gm_batch* b, double* out2)
{
    if(!b || !out2) return fail(GM_ERR_ARG, "null argument");
    out2[0] = b->plan->totalMs;
    out2[1] = sweepMs(b->plan);
    return GM_OK;
}

int gm_batch_kernel_times(const gm_batch* b, double* ms, int64_t* launches)
{
    if(!b) return fail(GM_ERR_ARG, "null batch");
    for(int k = 0; k < GM_KT_COUNT; ++k)
    {
        if(ms) ms[k] = b->plan->kernMs[k];
        if(launches) launches[k] = b->plan->kernLaunches[k];
    }
    return GM_OK;
}

int gm_batch_autotune(gm_batch* b, int smoother, int precision, int passes, int32_t* best, double* times)
{
    if(!b || !best) return fail(GM_ERR_ARG, "null argument");
    int rc = checkRun(smoother, passes, precision, GM_KERNEL_WAVE, 1.0);
    if(rc) return rc;
    if(passes < 1) return fail(GM_ERR_ARG, "autotuning needs at least one pass");
    if(smoother == gmd::SM_NELDER_MEAD) return fail(GM_ERR_ARG, "Nelder-Mead has no tunable launch shape");
    if((rc = checkBatchCrew(b))) return rc;
    GM_HIP(hipSetDevice(b->device));
    gmr::Plan* p = b->plan;
    // Every trial starts from the same positions; they are restored at the end.
    std::vector<std::vector<double4>> saved(p->meshes.size());
    for(size_t i = 0; i < p->meshes.size(); ++i)
    {
        saved[i].resize(p->meshes[i]->topo->nverts);
        if(!saved[i].empty())
            GM_HIP(hipMemcpyAsync(saved[i].data(), p->meshes[i]->dm.pos64, saved[i].size() * sizeof(double4),
                                  hipMemcpyDeviceToHost, p->stream));
    }
    GM_HIP(hipStreamSynchronize(p->stream));
    auto restore = [&]() -> int {
        for(size_t i = 0; i < p->meshes.size(); ++i)
            if(!saved[i].empty())
                GM_HIP(hipMemcpyAsync(p->meshes[i]->dm.pos64, saved[i].data(), saved[i].size() * sizeof(double4),
                                      hipMemcpyHostToDevice, p->stream));
        GM_HIP(hipStreamSynchronize(p->stream));
        return GM_OK;
    };
    const int PPG[4] = {1, 2, 4, 8};
    const int TILE[3] = {64, 128, 256};
    double bestMs = 1e300;
    gmd::LaneShape bestShape = b->lane[smoother].patchesPerGroup ? b->lane[smoother] : defaultLane(p->meshes, smoother);
    for(int a = 0; a < 4; ++a)
        for(int t = 0; t < 3; ++t)
        {
            const gmd::LaneShape shape{PPG[a], TILE[t]};
            // A shape is a candidate when the candidate-parallel kernel takes
            // every single-kind list with it; GETMe has no LDS tile, so only
            // its block size (patches per group) is swept.
            const bool getme = smoother == gmd::SM_GETME;
            bool fits = getme ? TILE[t] == 64 : true;
            for(const auto& col : p->colours)
                for(const gmr::VList& l : col)
                {
                    const bool f64 = precision == GM_FP64 || precision == GM_FP64_FAST;
                    gmd::VertexWork w{p->dTable, l.segs, l.nsegs, l.count, l.kind, l.maxPatch,
                                      precision == GM_FP64_FAST ? 1 : 0};
                    if(!getme && l.kind != gmd::EK_MIXED && !l.constrained &&
                       !gmd::laneFits(w, smoother, shape, f64, p->meshes[0]->evaluator, p->meshes[0]->sampler))
                        fits = false;
                }
            double ms = -1.0;
            if(fits)
            {
                gmr::RunParams r;
                r.smoother = smoother;
                r.passes = passes;
                r.precision = precision;
    r.polishPasses = gmr::polishPasses();
                r.kernel = GM_KERNEL_WAVE;
                r.spawnCoeff = 1.0 / 20.0;
                r.lane = shape;
                rc = gmr::runPlan(p, p->meshes[0]->crew, p->meshes[0]->evaluator, p->meshes[0]->sampler, r);
                if(rc) return rc;
                ms = sweepMs(p);
                if(ms < bestMs)
                {
                    bestMs = ms;
                    bestShape = shape;
                }
                if((rc = restore())) return rc;
            }
            if(times) times[a * 3 + t] = ms;
        }
    // FP32 runs leave pos32 stale relative to the restored pos64; every run
    // converts from pos64 first, so nothing else needs restoring.
    b->lane[smoother] = bestShape;
    best[0] = bestShape.patchesPerGroup;
    best[1] = bestShape.tileElems;
    return GM_OK;
}

int gm_throughput_probe(int op, double* rates)
{
    if(op < 0 || op > 2 || !rates) return fail(GM_ERR_ARG, "bad probe arguments");
    int ndev = 0, dev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GM_ERR_HIP, "no HIP device visible");
    GM_HIP(hipGetDevice(&dev));
    int cus = 0;
    GM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = 16 * cus, threads = 256, iters = 2048;
    float* out = nullptr;
    GM_HIP(devAlloc(&out, (size_t)blocks * threads));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if(e == hipSuccess) e = hipEventCreate(&e0);
    if(e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f, best = 0.0f;
    for(int rep = 0; rep < 4 && e == hipSuccess; ++rep)   // the first run warms up
    {
        e = hipEventRecord(e0, s);
        if(e == hipSuccess) e = gmd::launchThroughput(op, blocks, iters, out, s);
        if(e == hipSuccess) e = hipEventRecord(e1, s);
        if(e == hipSuccess) e = hipEventSynchronize(e1);
        if(e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if(e == hipSuccess && rep > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    if(e1) (void)hipEventDestroy(e1);
    if(e0) (void)hipEventDestroy(e0);
    if(s) (void)hipStreamDestroy(s);
    gmr::poolFree(out);
    if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("throughput probe: ") + hipGetErrorString(e));
    const double steps = (double)blocks * threads * iters * gmd::throughputChains() / (best * 1e-3);
    rates[0] = op == 1 ? 0.0 : 2.0 * steps;   // transcendental instructions per second (lane ops)
    rates[1] = op == 0 ? 0.0 : steps;         // FMA instructions per second (lane ops)
    rates[2] = best;
    return GM_OK;
}

int gm_math_probe(int op, const double* a, const double* b, double* out, int64_t n)
{
    if(op < 0 || op > 8 || n < 0 || (n && (!a || !out || ((op == 1 || op == 8) && !b))))
        return fail(GM_ERR_ARG, "bad probe arguments");
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
    gmr::poolFree(da);
    gmr::poolFree(db);
    gmr::poolFree(dout);
    return GM_OK;
}

} // extern "C"
