// noalloc.hip -- gm_evaluate_noalloc: the FP64 mesh-quality reduction of
// gm_mesh_evaluate over caller-owned device arrays, for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  It allocates nothing: the reduction's
// block partials, flags and histogram live in `scratch`, a static device
// array of fixed size that each GPU holds until the process ends.
//
// The element loop, the block decomposition (gmd::evalBlocksFor) and the
// reduction tree are those of evalKernel / evalFinalKernel, so the result has
// the bits of gm_mesh_evaluate on the same mesh: element qualities do not
// depend on vertex numbering, and elements are visited in the caller's order,
// which is the order the mesh handle keeps.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "gm_device.h"
#include "gm_reduce.h"
#include "gm_runtime.h"

namespace
{

using gmq::v3;
using gmq::mk;

struct NoAllocScratch
{
    double partials[3 * gmd::EVAL_MAX_BLOCKS];   // per block: min, sum of 1/q, compensation
    int32_t flags[gmd::EVAL_MAX_BLOCKS];         // per block: 1 when an element quality <= 0
    int32_t hist[20];
};

// The device array; host code reaches it only through the per-device
// address from hipGetSymbolAddress (the name itself is a host stand-in).
static __device__ __align__(256) NoAllocScratch scratch;

const int MAX_DEVICES = 64;
std::mutex g_mu;
NoAllocScratch* g_scratchOf[MAX_DEVICES] = {};

int scratchFor(int device, NoAllocScratch** out)
{
    if(device < 0 || device >= MAX_DEVICES) return gmr::fail(GM_ERR_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if(!g_scratchOf[device])
    {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if(e != hipSuccess || !p)
            return gmr::fail(GM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchOf[device] = (NoAllocScratch*)p;
    }
    *out = g_scratchOf[device];
    return GM_OK;
}

__device__ __forceinline__ v3<double> ldp(const double* pos, uint32_t v)
{
    return mk<double>(pos[3 * (size_t)v], pos[3 * (size_t)v + 1], pos[3 * (size_t)v + 2]);
}

__global__ void noallocClearKernel(NoAllocScratch* s)
{
    if(threadIdx.x < 20) s->hist[threadIdx.x] = 0;
}

// evalKernel's loop and block reduction over caller arrays: block b owns the
// contiguous element range [b chunk, (b + 1) chunk), tets before hexes.
template<int EV, int SA, bool HEX>
__global__ void __launch_bounds__(gmd::EVAL_THREADS) noallocEvalKernel(const double* __restrict__ pos,
                                                                      const uint32_t* __restrict__ tets,
                                                                      uint64_t ntets,
                                                                      const uint32_t* __restrict__ hexs,
                                                                      uint64_t nhexs, uint32_t blocks,
                                                                      gmq::Crew c, NoAllocScratch* s)
{
    __shared__ double smin[gmd::EVAL_THREADS], ssum[gmd::EVAL_THREADS], scmp[gmd::EVAL_THREADS];
    __shared__ int shist[20];
    __shared__ int sbad;
    if(threadIdx.x < 20) shist[threadIdx.x] = 0;
    if(threadIdx.x == 0) sbad = 0;
    __syncthreads();
    const uint64_t E = ntets + nhexs;
    const uint64_t chunk = (E + blocks - 1) / blocks;
    const uint64_t b0 = (uint64_t)blockIdx.x * chunk;
    const uint64_t b1 = (b0 + chunk < E) ? b0 + chunk : E;
    double mn = 1.0, sum = 0.0, cmp = 0.0;
    int bad = 0;
    for(uint64_t e = b0 + threadIdx.x; e < b1; e += blockDim.x)
    {
        double q;
        if(!HEX || e < ntets)
        {
            const uint32_t* t = tets + 4 * e;
            v3<double> vp[4] = {ldp(pos, t[0]), ldp(pos, t[1]), ldp(pos, t[2]), ldp(pos, t[3])};
            q = gmq::tetQuality<double, EV, SA>(c, vp);
        }
        else
        {
            const uint32_t* h = hexs + 8 * (e - ntets);
            v3<double> vp[8] = {ldp(pos, h[0]), ldp(pos, h[1]), ldp(pos, h[2]), ldp(pos, h[3]),
                                ldp(pos, h[4]), ldp(pos, h[5]), ldp(pos, h[6]), ldp(pos, h[7])};
            q = gmq::hexQuality<double, EV, SA>(c, vp);
        }
        mn = gmq::tmin(mn, q);
        if(q > 0.0)
        {
            double t, err;
            gmd::twoSum(sum, 1.0 / q, t, err);
            sum = t;
            cmp += err;
        }
        else
            bad = 1;
        const int bucket = gmd::histBucket(q);
        unsigned long long pending = __ballot(true);
        while(pending)
        {
            const int leader = (int)__builtin_ctzll(pending);
            const int lb = __builtin_amdgcn_readlane(bucket, leader);
            const unsigned long long same = __ballot(bucket == lb);
            if((int)(threadIdx.x & 63) == leader)
                atomicAdd(&shist[lb], (int)__builtin_popcountll(same));
            pending &= ~same;
        }
    }
    smin[threadIdx.x] = mn;
    ssum[threadIdx.x] = sum;
    scmp[threadIdx.x] = cmp;
    if(bad) sbad = 1;
    __syncthreads();
    for(int w = gmd::EVAL_THREADS / 2; w > 0; w >>= 1)
    {
        if(threadIdx.x < (unsigned)w)
        {
            smin[threadIdx.x] = gmq::tmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            double t, err;
            gmd::twoSum(ssum[threadIdx.x], ssum[threadIdx.x + w], t, err);
            ssum[threadIdx.x] = t;
            scmp[threadIdx.x] = scmp[threadIdx.x] + scmp[threadIdx.x + w] + err;
        }
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        s->partials[3 * blockIdx.x] = smin[0];
        s->partials[3 * blockIdx.x + 1] = ssum[0];
        s->partials[3 * blockIdx.x + 2] = scmp[0];
        s->flags[blockIdx.x] = sbad;
    }
    if(threadIdx.x < 20 && shist[threadIdx.x])
        atomicAdd(&s->hist[threadIdx.x], shist[threadIdx.x]);
}

// evalFinalKernel's reduction of the block partials; writes the caller's
// outputs {min, sum of 1/q (+inf if any q <= 0), n} and the histogram.
__global__ void __launch_bounds__(gmd::EVAL_THREADS) noallocFinalKernel(const NoAllocScratch* __restrict__ s,
                                                                       uint32_t blocks, double n, double* out3,
                                                                       int32_t* hist20)
{
    __shared__ double smin[gmd::EVAL_THREADS], ssum[gmd::EVAL_THREADS], scmp[gmd::EVAL_THREADS];
    __shared__ int sbad;
    if(threadIdx.x == 0) sbad = 0;
    __syncthreads();
    double mn = 1.0, sum = 0.0, cmp = 0.0;
    for(uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
    {
        mn = gmq::tmin(mn, s->partials[3 * b]);
        double t, err;
        gmd::twoSum(sum, s->partials[3 * b + 1], t, err);
        sum = t;
        cmp += err + s->partials[3 * b + 2];
        if(s->flags[b]) sbad = 1;
    }
    smin[threadIdx.x] = mn;
    ssum[threadIdx.x] = sum;
    scmp[threadIdx.x] = cmp;
    __syncthreads();
    for(int w = gmd::EVAL_THREADS / 2; w > 0; w >>= 1)
    {
        if(threadIdx.x < (unsigned)w)
        {
            smin[threadIdx.x] = gmq::tmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            double t, err;
            gmd::twoSum(ssum[threadIdx.x], ssum[threadIdx.x + w], t, err);
            ssum[threadIdx.x] = t;
            scmp[threadIdx.x] = scmp[threadIdx.x] + scmp[threadIdx.x + w] + err;
        }
        __syncthreads();
    }
    if(threadIdx.x == 0)
    {
        out3[0] = smin[0];
        out3[1] = sbad ? (double)__builtin_huge_val() : ssum[0] + scmp[0];
        out3[2] = n;
    }
    if(threadIdx.x < 20) hist20[threadIdx.x] = s->hist[threadIdx.x];
}

template<int EV, int SA>
hipError_t launchNoAlloc(const double* pos, const uint32_t* tets, uint64_t ntets, const uint32_t* hexs,
                         uint64_t nhexs, uint32_t blocks, const gmq::Crew& c, NoAllocScratch* s, hipStream_t stream)
{
    if(nhexs)
        noallocEvalKernel<EV, SA, true><<<blocks, gmd::EVAL_THREADS, 0, stream>>>(pos, tets, ntets, hexs, nhexs,
                                                                                 blocks, c, s);
    else
        noallocEvalKernel<EV, SA, false><<<blocks, gmd::EVAL_THREADS, 0, stream>>>(pos, tets, ntets, hexs, nhexs,
                                                                                  blocks, c, s);
    return hipGetLastError();
}

} // namespace

extern "C" int gm_evaluate_noalloc(const double* d_pos, int64_t nverts, const uint32_t* d_tets, int64_t ntets,
                                   const uint32_t* d_hexs, int64_t nhexs, int evaluator, int sampler, double scaling,
                                   double* d_out3, int32_t* d_hist20, void* stream)
{
    using gmr::fail;
    if(nverts < 0 || ntets < 0 || nhexs < 0) return fail(GM_ERR_ARG, "negative size");
    if(nverts > GM_NOALLOC_MAX_VERTS)
        return fail(GM_ERR_ARG, "gm_evaluate_noalloc takes at most GM_NOALLOC_MAX_VERTS (2^32 - 1) vertices");
    if(ntets + nhexs > GM_NOALLOC_MAX_ELEMS)
        return fail(GM_ERR_ARG, "gm_evaluate_noalloc takes at most GM_NOALLOC_MAX_ELEMS (2^31) elements");
    if(evaluator < 0 || evaluator > 1) return fail(GM_ERR_ARG, "unknown evaluator");
    if(sampler != GM_UNIFORM && sampler != GM_ANALYTIC_SHELL)
        return fail(GM_ERR_ARG, "gm_evaluate_noalloc takes the uniform or analytic sampler (the texture sampler "
                                "needs its texels on the device)");
    if(!(scaling > 0.0)) return fail(GM_ERR_ARG, "scaling must be positive");
    if(!d_out3 || !d_hist20) return fail(GM_ERR_ARG, "null output");
    if((nverts && !d_pos) || (ntets && !d_tets) || (nhexs && !d_hexs)) return fail(GM_ERR_ARG, "null array");
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    NoAllocScratch* s = nullptr;
    int rc = scratchFor(device, &s);
    if(rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    gmq::Crew c{};
    c.scaling = scaling;
    c.scaling2 = scaling * scaling;
    c.scaling3 = c.scaling2 * scaling;
    const uint64_t E = (uint64_t)ntets + (uint64_t)nhexs;
    const uint32_t blocks = gmd::evalBlocksFor(E);
    // The histogram accumulates across blocks: it starts at zero on every
    // call (a fill kernel, so the call can be captured in a graph).
    noallocClearKernel<<<1, 64, 0, st>>>(s);
    e = hipGetLastError();
    if(e == hipSuccess)
    {
        const int combo = evaluator * 2 + (sampler == GM_ANALYTIC_SHELL ? 1 : 0);
        if(E)
            e = combo == 0 ? launchNoAlloc<0, 0>(d_pos, d_tets, ntets, d_hexs, nhexs, blocks, c, s, st)
              : combo == 1 ? launchNoAlloc<0, 1>(d_pos, d_tets, ntets, d_hexs, nhexs, blocks, c, s, st)
              : combo == 2 ? launchNoAlloc<1, 0>(d_pos, d_tets, ntets, d_hexs, nhexs, blocks, c, s, st)
                           : launchNoAlloc<1, 1>(d_pos, d_tets, ntets, d_hexs, nhexs, blocks, c, s, st);
    }
    if(e == hipSuccess)
    {
        noallocFinalKernel<<<1, gmd::EVAL_THREADS, 0, st>>>(s, E ? blocks : 0, (double)E, d_out3, d_hist20);
        e = hipGetLastError();
    }
    if(e != hipSuccess) return fail(GM_ERR_HIP, std::string("gm_evaluate_noalloc launch: ") + hipGetErrorString(e));
    return GM_OK;
}
