// server.hip -- the serving loop: smoothing jobs for meshes of uneven size
// are admitted by free device memory, stay resident until their last pass,
// and advance together one pass per round (continuous batching).
//
// Threads:
//  * workers: build a submitted job's host topology through the shared
//    topology cache and compute its device footprint; once the scheduler has
//    admitted the job, upload it and start it (pass record 0) on the job's
//    own stream; once its last pass ran, copy its results out and release its
//    device memory.  Retirements go first, then uploads, then builds, so the
//    memory of finished jobs comes back before new work is taken on.
//  * one scheduler: admits built jobs in submission order against the memory
//    budget and the slot count, and runs rounds.  A round launches one pass
//    for every running job -- jobs grouped by (smoother, FP64 pass, widening
//    first, polish pass), each group one fused plan on its own stream -- and
//    waits for it.  Jobs the workers started meanwhile join the next round,
//    so uploads and result copies overlap the rounds' kernels.
//
// A job's kernels, their order within the job and its reduction trees are the
// ones gm_mesh_smooth runs for the same mesh alone, so every result is
// bit-identical to smoothing that mesh alone with the same precision mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gm_runtime.h"

using gmr::fail;
using Clock = std::chrono::steady_clock;

namespace
{

// Public states (gm_server_wait): queued, built, active, done, failed.
enum JobState { JOB_QUEUED = 0, JOB_BUILT = 1, JOB_ACTIVE = 2, JOB_DONE = 3, JOB_FAILED = 4 };

struct Job
{
    int64_t id = 0;
    // Inputs: owned copies, or the caller's arrays when borrowed; dropped
    // once the topology is built.
    int64_t nverts = 0, ntets = 0, nhexs = 0;
    std::vector<double> posCopy;
    std::vector<int32_t> fixedCopy;
    std::vector<uint32_t> tetsCopy, hexsCopy;
    const double* pos = nullptr;
    const int32_t* fixed = nullptr;
    const uint32_t* tets = nullptr;
    const uint32_t* hexs = nullptr;
    int smoother = 0, passes = 0, precision = 0;
    int polishPasses = 1;   // gm_set_polish_passes at submission
    // State.
    int state = JOB_QUEUED;
    gm_mesh* mesh = nullptr;
    size_t bytes = 0;
    int pass = 0;
    bool inF32 = false;
    // Results.
    std::vector<double> outPos, stats;
    std::vector<int32_t> hist;
    int errCode = 0;
    std::string err;
    Clock::time_point tSubmit, tBuilt, tAdmit, tDone;

    void dropInputs()
    {
        std::vector<double>().swap(posCopy);
        std::vector<int32_t>().swap(fixedCopy);
        std::vector<uint32_t>().swap(tetsCopy);
        std::vector<uint32_t>().swap(hexsCopy);
        pos = nullptr;
        fixed = nullptr;
        tets = nullptr;
        hexs = nullptr;
    }
};

double ms(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

} // namespace

struct gm_server
{
    int device = 0;
    size_t budget = 0;
    int maxActive = 0;
    int evaluator = 0, sampler = 0;
    gmq::Crew crew{};
    float* texels = nullptr;         // the texture grid every job samples (GM_TEXTURE)
    gmd::LaneShape lane[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};

    std::mutex mu;
    std::condition_variable cvSched, cvWork, cvDone;
    bool stop = false;
    int64_t nextId = 1;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job*> buildQueue;     // submitted, topology not built
    std::deque<Job*> pending;        // submitted, not admitted, in submission order
    std::deque<Job*> uploadQueue;    // admitted, to upload and start
    std::deque<Job*> retireQueue;    // last pass run, results to copy out
    std::vector<Job*> ready;         // started, waiting for the next round
    size_t reserved = 0;             // device bytes of the admitted jobs
    int admitted = 0;                // jobs holding a slot: uploading, ready, running or retiring

    // Scheduler-owned.
    std::vector<Job*> active;
    struct Group
    {
        std::vector<int64_t> ids;
        gmr::Plan* plan = nullptr;
        hipStream_t stream = nullptr;
    };
    std::map<std::tuple<int, int, int, int>, Group> groups;

    // Statistics (under mu).
    int64_t submitted = 0, done = 0, failed = 0, rounds = 0, peakActive = 0;
    size_t peakReserved = 0;
    double roundMs = 0.0;

    std::vector<std::thread> workers;
    std::thread scheduler;
};

namespace
{

// Both with s->mu held.
void finishJob(gm_server* s, Job* j, int code, const std::string& msg)
{
    j->errCode = code;
    j->err = msg;
    j->state = code ? JOB_FAILED : JOB_DONE;
    j->tDone = Clock::now();
    if(code) ++s->failed;
    else ++s->done;
    s->cvDone.notify_all();
}

void releaseSlot(gm_server* s, Job* j)
{
    s->reserved -= j->bytes;
    --s->admitted;
    s->cvSched.notify_all();
}

// ---- worker tasks, run without the lock

void buildJob(gm_server* s, Job* j)
{
    gm_mesh* m = nullptr;
    int rc = gmr::makeMesh(j->nverts, j->pos, j->fixed, j->ntets, j->tets, j->nhexs, j->hexs, &m);
    std::string msg = rc ? gmr::lastError() : "";
    size_t bytes = 0;
    if(!rc)
    {
        const gmt::Topology& t = *m->topo;
        for(int64_t v = 0; v < t.nverts && !rc; ++v)
            if(!t.fixed[v] && t.eoff[v + 1] == t.eoff[v])
            {
                rc = GM_ERR_ARG;
                msg = "a free vertex belongs to no element";
            }
        bytes = gmr::residentBytes(m, j->passes + 1, j->smoother == gmd::SM_GETME);
    }
    std::lock_guard<std::mutex> lock(s->mu);
    j->dropInputs();
    j->tBuilt = Clock::now();
    if(!rc && bytes > s->budget)
    {
        rc = GM_ERR_NOMEM;
        msg = "the job needs more device memory than the server's budget";
    }
    if(rc)
    {
        if(m) gmr::destroyMesh(m);
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), j));
        finishJob(s, j, rc, msg);
    }
    else
    {
        j->mesh = m;
        j->bytes = bytes;
        j->state = JOB_BUILT;
    }
    s->cvSched.notify_all();
}

// Uploads an admitted job, sets its crew and runs its start (counters, FP32
// copy, pass record 0) on the job's own stream.
void uploadJob(gm_server* s, Job* j)
{
    gm_mesh* m = j->mesh;
    int rc = gmr::uploadMesh(m);
    if(!rc)
    {
        m->crew = s->crew;
        m->evaluator = s->evaluator;
        m->sampler = s->sampler;
        rc = gmr::preparePlan(m->plan, j->passes + 1, j->smoother == gmd::SM_GETME);
    }
    j->inF32 = gmr::startsF32(j->precision);
    j->pass = 0;
    if(!rc) rc = gmr::launchBegin(m->plan, s->crew, s->evaluator, s->sampler, j->inF32);
    if(!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = fail(GM_ERR_HIP, "job start failed");
    const std::string msg = rc ? gmr::lastError() : "";
    if(rc)
    {
        gmr::destroyMesh(m);
        j->mesh = nullptr;
    }
    std::lock_guard<std::mutex> lock(s->mu);
    if(rc)
    {
        releaseSlot(s, j);
        finishJob(s, j, rc, msg);
    }
    else
        s->ready.push_back(j);
    s->cvSched.notify_all();
}

// Copies a finished job's results (pass records, positions in the caller's
// numbering) and releases its device memory.
void retireJob(gm_server* s, Job* j)
{
    gm_mesh* m = j->mesh;
    int rc = GM_OK;
    if(j->inF32)
    {
        rc = gmr::launchToF64All(m->plan);
        j->inF32 = false;
    }
    const int records = j->passes + 1;
    j->stats.assign(3 * (size_t)records, 0.0);
    j->hist.assign(20 * (size_t)records, 0);
    if(!rc) rc = gmr::fetchRecords(m->plan, 0, records, j->stats.data(), j->hist.data(), nullptr);
    const gmt::Topology& t = *m->topo;
    const size_t V = (size_t)t.nverts;
    if(!rc)
    {
        std::vector<double4> p(V);
        hipError_t e = V ? hipMemcpyAsync(p.data(), m->dm.pos64, V * sizeof(double4), hipMemcpyDeviceToHost,
                                          m->stream)
                         : hipSuccess;
        if(e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if(e != hipSuccess) rc = fail(GM_ERR_HIP, std::string("result copy: ") + hipGetErrorString(e));
        j->outPos.resize(3 * V);
        for(size_t old = 0; old < V && !rc; ++old)
        {
            const double4 q = p[m->slotOf[t.permutation[old]]];
            j->outPos[3 * old] = q.x;
            j->outPos[3 * old + 1] = q.y;
            j->outPos[3 * old + 2] = q.z;
        }
    }
    const std::string msg = rc ? gmr::lastError() : "";
    gmr::destroyMesh(m);
    j->mesh = nullptr;
    std::lock_guard<std::mutex> lock(s->mu);
    releaseSlot(s, j);
    finishJob(s, j, rc, msg);
}

void workerLoop(gm_server* s)
{
    (void)hipSetDevice(s->device);
    for(;;)
    {
        Job* j = nullptr;
        int task = 0;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->cvWork.wait(lock, [&] {
                return s->stop || !s->retireQueue.empty() || !s->uploadQueue.empty() || !s->buildQueue.empty();
            });
            if(s->stop) return;
            std::deque<Job*>& q = !s->retireQueue.empty() ? s->retireQueue
                                : !s->uploadQueue.empty() ? s->uploadQueue
                                                          : s->buildQueue;
            task = &q == &s->retireQueue ? 2 : &q == &s->uploadQueue ? 1 : 0;
            j = q.front();
            q.pop_front();
        }
        if(task == 2) retireJob(s, j);
        else if(task == 1) uploadJob(s, j);
        else buildJob(s, j);
    }
}

// ---- admission (s->mu held): pending jobs in submission order; a built job
// that does not fit the memory left blocks those behind it (no overtaking of
// large jobs), a job still building is skipped.

Job* nextAdmissible(gm_server* s)
{
    if(s->admitted >= s->maxActive) return nullptr;
    for(Job* j : s->pending)
        if(j->state == JOB_BUILT) return s->reserved + j->bytes <= s->budget ? j : nullptr;
    return nullptr;
}

void admit(gm_server* s)
{
    while(Job* j = nextAdmissible(s))
    {
        s->pending.erase(std::find(s->pending.begin(), s->pending.end(), j));
        s->reserved += j->bytes;
        s->peakReserved = std::max(s->peakReserved, s->reserved);
        ++s->admitted;
        j->state = JOB_ACTIVE;
        j->tAdmit = Clock::now();
        s->uploadQueue.push_back(j);
        s->cvWork.notify_one();
    }
}

// Brings the group plans in line with the active set and launches one pass
// per group.
int launchRound(gm_server* s, const gmd::LaneShape* lanes)
{
    // Groups: (smoother, FP64 pass, widening first, polish pass).  A polish
    // pass runs the mixed-precision FP64 kernels, an exact-mode pass the
    // bit-exact ones, as gm_mesh_smooth does.
    std::map<std::tuple<int, int, int, int>, std::vector<Job*>> members;
    for(Job* j : s->active)
    {
        const bool f64 = gmr::passF64(j->precision, j->pass, j->passes, j->polishPasses);
        const bool polish = f64 && gmr::passFast(j->precision);
        const bool convert = f64 && j->inF32;
        members[std::make_tuple(j->smoother, (int)f64, (int)convert, (int)polish)].push_back(j);
    }
    // Drop plans whose membership changed.
    for(auto& kv : s->groups)
    {
        auto it = members.find(kv.first);
        std::vector<int64_t> ids;
        if(it != members.end())
            for(Job* j : it->second) ids.push_back(j->id);
        if(ids != kv.second.ids && kv.second.plan)
        {
            gmr::freePlan(kv.second.plan);
            kv.second.plan = nullptr;
            kv.second.ids.clear();
        }
    }
    for(auto& kv : members)
    {
        gm_server::Group& g = s->groups[kv.first];
        if(!g.stream && hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess)
            return fail(GM_ERR_HIP, "stream creation failed");
        if(!g.plan)
        {
            std::vector<gm_mesh*> ms;
            for(Job* j : kv.second)
            {
                ms.push_back(j->mesh);
                g.ids.push_back(j->id);
            }
            int rc = gmr::buildPlan(ms, g.stream, &g.plan);
            if(!rc) rc = gmr::preparePlan(g.plan, 0, false);
            if(rc)
            {
                g.ids.clear();
                return rc;
            }
        }
        const int smoother = std::get<0>(kv.first);
        gmr::RunParams r;
        r.smoother = smoother;
        r.precision = std::get<3>(kv.first) ? GM_FP32_POLISH : (std::get<1>(kv.first) ? GM_FP64 : GM_FP32);
        r.kernel = GM_KERNEL_WAVE;
        r.spawnCoeff = 1.0 / 20.0;
        if(lanes[smoother].patchesPerGroup)
            r.lane = lanes[smoother];
        else
            r.lane = gmd::LaneShape{smoother == gmd::SM_GETME ? 4 : 1, 0};   // tile per work list
        const int rc = gmr::launchPass(g.plan, s->crew, s->evaluator, s->sampler, r, std::get<1>(kv.first) != 0,
                                       std::get<2>(kv.first) != 0);
        if(rc) return rc;
    }
    return GM_OK;
}

void schedulerLoop(gm_server* s)
{
    (void)hipSetDevice(s->device);
    for(;;)
    {
        gmd::LaneShape lanes[4];
        {
            std::unique_lock<std::mutex> lock(s->mu);
            s->cvSched.wait(lock, [&] {
                return s->stop || !s->active.empty() || !s->ready.empty() || nextAdmissible(s) != nullptr;
            });
            if(s->stop) break;
            admit(s);
            // Started jobs join this round; a job with no passes retires at once.
            for(Job* j : s->ready)
            {
                if(j->passes > 0) s->active.push_back(j);
                else
                {
                    s->retireQueue.push_back(j);
                    s->cvWork.notify_one();
                }
            }
            s->ready.clear();
            for(int k = 0; k < 4; ++k) lanes[k] = s->lane[k];
            s->peakActive = std::max<int64_t>(s->peakActive, (int64_t)s->active.size());
        }
        if(s->active.empty()) continue;
        const Clock::time_point t0 = Clock::now();
        const int rc = launchRound(s, lanes);
        // Also after a failed launch: whatever was launched must end before
        // the memory goes back.
        for(auto& kv : s->groups)
            if(kv.second.stream) (void)hipStreamSynchronize(kv.second.stream);
        const Clock::time_point t1 = Clock::now();
        std::lock_guard<std::mutex> lock(s->mu);
        if(rc)
        {
            // A launch failure fails every job of the round.
            const std::string msg = gmr::lastError();
            for(Job* j : s->active)
            {
                gmr::destroyMesh(j->mesh);
                j->mesh = nullptr;
                releaseSlot(s, j);
                finishJob(s, j, rc, msg);
            }
            s->active.clear();
            continue;
        }
        std::vector<Job*> still;
        for(Job* j : s->active)
        {
            const bool f64 = gmr::passF64(j->precision, j->pass, j->passes, j->polishPasses);
            if(f64) j->inF32 = false;
            ++j->pass;
            if(j->pass < j->passes) still.push_back(j);
            else
            {
                s->retireQueue.push_back(j);
                s->cvWork.notify_one();
            }
        }
        s->active.swap(still);
        ++s->rounds;
        s->roundMs += ms(t0, t1);
    }
    // Shutdown: cancel the jobs the scheduler holds; the device memory of
    // jobs left in the worker queues is released by gm_server_destroy.
    std::vector<Job*> held;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        held.swap(s->active);
        held.insert(held.end(), s->ready.begin(), s->ready.end());
        s->ready.clear();
    }
    for(Job* j : held)
    {
        gmr::destroyMesh(j->mesh);
        j->mesh = nullptr;
        std::lock_guard<std::mutex> lock(s->mu);
        releaseSlot(s, j);
        finishJob(s, j, GM_ERR_ARG, "server shut down");
    }
    for(auto& kv : s->groups)
    {
        gmr::freePlan(kv.second.plan);
        if(kv.second.stream) (void)hipStreamDestroy(kv.second.stream);
    }
    s->groups.clear();
}

int checkRequest(int smoother, int passes, int precision)
{
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(passes < 0) return fail(GM_ERR_ARG, "negative pass count");
    if(precision < 0 || precision > 3) return fail(GM_ERR_ARG, "unknown precision");
    return GM_OK;
}

// Queues a job whose inputs are set (s->mu not held).
int enqueue(gm_server* s, std::unique_ptr<Job> j, int64_t* jobId)
{
    std::lock_guard<std::mutex> lock(s->mu);
    if(s->stop) return fail(GM_ERR_ARG, "server is shutting down");
    if(s->sampler == GM_TEXTURE && !s->crew.texels)
        return fail(GM_ERR_ARG, "texture sampler: gm_server_set_texture must come before the first job");
    j->id = s->nextId++;
    *jobId = j->id;
    Job* raw = j.get();
    s->jobs.emplace(raw->id, std::move(j));
    s->buildQueue.push_back(raw);
    s->pending.push_back(raw);
    ++s->submitted;
    s->cvWork.notify_one();
    return GM_OK;
}

} // namespace

extern "C" {

int gm_server_create(int64_t budgetBytes, int maxActive, int workerThreads, int evaluator, int sampler,
                     double scaling, gm_server** out)
{
    if(!out) return fail(GM_ERR_ARG, "out is null");
    *out = nullptr;
    if(evaluator < 0 || evaluator > 1) return fail(GM_ERR_ARG, "unknown evaluator");
    if(sampler < 0 || sampler > 3) return fail(GM_ERR_ARG, "unknown sampler");
    if(!(scaling > 0.0)) return fail(GM_ERR_ARG, "scaling must be positive");
    if(maxActive < 1 || workerThreads < 1 || workerThreads > 64) return fail(GM_ERR_ARG, "bad slot or thread count");
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GM_ERR_HIP, "no HIP device visible");
    auto s = std::make_unique<gm_server>();
    if(hipGetDevice(&s->device) != hipSuccess) return fail(GM_ERR_HIP, "hipGetDevice failed");
    size_t freeB = 0, totalB = 0;
    if(hipMemGetInfo(&freeB, &totalB) != hipSuccess) return fail(GM_ERR_HIP, "hipMemGetInfo failed");
    // The budget never exceeds what is free now, less 1 GiB for plans,
    // streams and the runtime.
    const size_t headroom = (size_t)1 << 30;
    const size_t avail = freeB > headroom ? freeB - headroom : 0;
    s->budget = budgetBytes > 0 ? std::min((size_t)budgetBytes, avail) : avail;
    s->maxActive = maxActive;
    s->evaluator = evaluator;
    s->sampler = sampler;
    s->crew.scaling = scaling;
    s->crew.scaling2 = scaling * scaling;
    s->crew.scaling3 = s->crew.scaling2 * scaling;
    gmr::setAnalyticField(s->crew, 1.0);
    gm_server* raw = s.get();
    for(int i = 0; i < workerThreads; ++i) raw->workers.emplace_back(workerLoop, raw);
    raw->scheduler = std::thread(schedulerLoop, raw);
    *out = s.release();
    return GM_OK;
}

int gm_server_destroy(gm_server* s)
{
    if(!s) return GM_OK;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cvWork.notify_all();
    s->cvSched.notify_all();
    for(std::thread& t : s->workers) t.join();
    s->scheduler.join();
    // Jobs left in the worker queues (or started after the scheduler ended)
    // still hold a host topology or device memory.
    for(auto& kv : s->jobs)
        if(kv.second->mesh) gmr::destroyMesh(kv.second->mesh);
    gmr::poolFree(s->texels);
    delete s;
    return GM_OK;
}

int gm_server_set_texture(gm_server* s, const int32_t* tsize, const double* tmin, const double* text,
                          const float* texels)
{
    if(!s) return fail(GM_ERR_ARG, "null server");
    if(s->sampler != GM_TEXTURE) return fail(GM_ERR_ARG, "the server was not created with the texture sampler");
    if(!tsize || !tmin || !text || !texels) return fail(GM_ERR_ARG, "texture sampler needs a grid");
    gmq::Crew c = s->crew;
    for(int k = 0; k < 3; ++k)
    {
        if(tsize[k] < 1) return fail(GM_ERR_ARG, "texture size must be >= 1");
        if(!(text[k] > 0.0)) return fail(GM_ERR_ARG, "texture extents must be positive");
        c.tsize[k] = tsize[k];
        c.tmin[k] = tmin[k];
        c.text[k] = text[k];
    }
    gmq::textureFields(c);
    std::lock_guard<std::mutex> lock(s->mu);
    // Jobs read the crew without the lock, so it is fixed once one is queued.
    if(s->submitted) return fail(GM_ERR_ARG, "the texture is set before the first job is submitted");
    if(hipSetDevice(s->device) != hipSuccess) return fail(GM_ERR_HIP, "hipSetDevice failed");
    const size_t n = 6 * (size_t)tsize[0] * tsize[1] * tsize[2];
    float* fresh = nullptr;
    hipError_t e = gmr::poolAlloc((void**)&fresh, n * sizeof(float));
    if(e != hipSuccess) return fail(e == hipErrorOutOfMemory ? GM_ERR_NOMEM : GM_ERR_HIP, "texture allocation failed");
    e = hipMemcpy(fresh, texels, n * sizeof(float), hipMemcpyHostToDevice);
    if(e != hipSuccess)
    {
        gmr::poolFree(fresh);
        return fail(GM_ERR_HIP, std::string("texture upload: ") + hipGetErrorString(e));
    }
    gmr::poolFree(s->texels);
    s->texels = fresh;
    c.texels = fresh;
    s->crew = c;
    return GM_OK;
}

int gm_server_set_shape(gm_server* s, int smoother, int patchesPerGroup, int tileElems)
{
    if(!s) return fail(GM_ERR_ARG, "null server");
    if(smoother < 0 || smoother > 3) return fail(GM_ERR_ARG, "unknown smoother");
    if(patchesPerGroup != 1 && patchesPerGroup != 2 && patchesPerGroup != 4 && patchesPerGroup != 8)
        return fail(GM_ERR_ARG, "patches per workgroup must be 1, 2, 4 or 8");
    if(tileElems != 64 && tileElems != 128 && tileElems != 256)
        return fail(GM_ERR_ARG, "LDS tile must be 64, 128 or 256 elements");
    std::lock_guard<std::mutex> lock(s->mu);
    s->lane[smoother] = gmd::LaneShape{patchesPerGroup, tileElems};
    return GM_OK;
}


int gm_server_submit_job(gm_server* s, const void* blob, int64_t bytes, int64_t* jobId)
{
    const Clock::time_point now = Clock::now();
    if(!s || !jobId) return fail(GM_ERR_ARG, "null argument");
    int64_t counts[3];
    int32_t request[3];
    int rc = gm_job_info(blob, bytes, counts, request);
    if(rc) return rc;
    if((rc = checkRequest(request[0], request[1], request[2]))) return rc;
    auto j = std::make_unique<Job>();
    j->nverts = counts[0];
    j->ntets = counts[1];
    j->nhexs = counts[2];
    j->posCopy.resize(3 * (size_t)j->nverts);
    j->fixedCopy.resize((size_t)j->nverts);
    j->tetsCopy.resize(4 * (size_t)j->ntets);
    j->hexsCopy.resize(8 * (size_t)j->nhexs);
    if((rc = gm_job_decode(blob, bytes, j->posCopy.data(), j->fixedCopy.data(), j->tetsCopy.data(),
                           j->hexsCopy.data())))
        return rc;
    j->pos = j->posCopy.data();
    j->fixed = j->fixedCopy.data();
    j->tets = j->tetsCopy.data();
    j->hexs = j->hexsCopy.data();
    j->smoother = request[0];
    j->passes = request[1];
    j->precision = request[2];
    j->polishPasses = gmr::polishPasses();
    j->tSubmit = now;
    return enqueue(s, std::move(j), jobId);
}

int gm_server_submit(gm_server* s, int64_t nverts, const double* pos, const int32_t* fixed, int64_t ntets,
                     const uint32_t* tets, int64_t nhexs, const uint32_t* hexs, int smoother, int passes,
                     int precision, int borrow, int64_t* jobId)
{
    const Clock::time_point now = Clock::now();
    if(!s || !jobId) return fail(GM_ERR_ARG, "null argument");
    if(nverts < 0 || ntets < 0 || nhexs < 0) return fail(GM_ERR_ARG, "negative size");
    if((nverts && (!pos || !fixed)) || (ntets && !tets) || (nhexs && !hexs)) return fail(GM_ERR_ARG, "null array");
    int rc = checkRequest(smoother, passes, precision);
    if(rc) return rc;
    auto j = std::make_unique<Job>();
    j->nverts = nverts;
    j->ntets = ntets;
    j->nhexs = nhexs;
    if(borrow)
    {
        j->pos = pos;
        j->fixed = fixed;
        j->tets = tets;
        j->hexs = hexs;
    }
    else
    {
        if(nverts)
        {
            j->posCopy.assign(pos, pos + 3 * (size_t)nverts);
            j->fixedCopy.assign(fixed, fixed + (size_t)nverts);
        }
        if(ntets) j->tetsCopy.assign(tets, tets + 4 * (size_t)ntets);
        if(nhexs) j->hexsCopy.assign(hexs, hexs + 8 * (size_t)nhexs);
        j->pos = j->posCopy.data();
        j->fixed = j->fixedCopy.data();
        j->tets = j->tetsCopy.data();
        j->hexs = j->hexsCopy.data();
    }
    j->smoother = smoother;
    j->passes = passes;
    j->precision = precision;
    j->polishPasses = gmr::polishPasses();
    j->tSubmit = now;
    return enqueue(s, std::move(j), jobId);
}

int gm_server_wait(gm_server* s, int64_t jobId, double timeoutMs, int* state)
{
    if(!s) return fail(GM_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->jobs.find(jobId);
    if(it == s->jobs.end()) return fail(GM_ERR_ARG, "unknown job");
    Job* j = it->second.get();
    auto finished = [&] { return j->state == JOB_DONE || j->state == JOB_FAILED; };
    if(timeoutMs < 0)
        s->cvDone.wait(lock, finished);
    else
        s->cvDone.wait_for(lock, std::chrono::duration<double, std::milli>(timeoutMs), finished);
    if(state) *state = j->state;
    return GM_OK;
}

int gm_server_result(gm_server* s, int64_t jobId, double* pos, double* stats, int32_t* hist, double* times4)
{
    if(!s) return fail(GM_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(jobId);
    if(it == s->jobs.end()) return fail(GM_ERR_ARG, "unknown job");
    Job* j = it->second.get();
    if(j->state == JOB_FAILED) return fail(j->errCode, "job failed: " + j->err);
    if(j->state != JOB_DONE) return fail(GM_ERR_ARG, "job has not finished");
    if(pos && !j->outPos.empty()) std::memcpy(pos, j->outPos.data(), j->outPos.size() * sizeof(double));
    if(stats) std::memcpy(stats, j->stats.data(), j->stats.size() * sizeof(double));
    if(hist) std::memcpy(hist, j->hist.data(), j->hist.size() * sizeof(int32_t));
    if(times4)
    {
        times4[0] = ms(j->tSubmit, j->tBuilt);
        times4[1] = ms(j->tBuilt, j->tAdmit);
        times4[2] = ms(j->tAdmit, j->tDone);
        times4[3] = ms(j->tSubmit, j->tDone);
    }
    return GM_OK;
}

int gm_server_release(gm_server* s, int64_t jobId)
{
    if(!s) return fail(GM_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->jobs.find(jobId);
    if(it == s->jobs.end()) return fail(GM_ERR_ARG, "unknown job");
    if(it->second->state != JOB_DONE && it->second->state != JOB_FAILED)
        return fail(GM_ERR_ARG, "job has not finished");
    s->jobs.erase(it);
    return GM_OK;
}

int gm_server_stats(gm_server* s, double* out8)
{
    if(!s || !out8) return fail(GM_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    out8[0] = (double)s->submitted;
    out8[1] = (double)s->done;
    out8[2] = (double)s->failed;
    out8[3] = (double)s->rounds;
    out8[4] = (double)s->peakActive;
    out8[5] = (double)s->peakReserved;
    out8[6] = (double)s->budget;
    out8[7] = s->roundMs;
    return GM_OK;
}

} // extern "C"
