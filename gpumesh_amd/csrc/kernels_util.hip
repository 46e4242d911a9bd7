// kernels_util.hip -- precision conversions, run resets and the device math
// probe.  Built with -ffp-contract=off like every FP64 kernel.
#include "gm_kernels.h"

namespace gmd
{

__global__ void __launch_bounds__(VERT_BLOCK) toF32Kernel(const DevMesh* __restrict__ meshes,
                                                  const uint4* __restrict__ spans, uint32_t nspans)
{
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    uint32_t i = bl.y * VERT_BLOCK + threadIdx.x;
    if(i >= m.nverts) return;
    double4 a = m.pos64[i];
    m.pos32[i] = make_float4((float)a.x, (float)a.y, (float)a.z, 0.0f);
}

__global__ void __launch_bounds__(VERT_BLOCK) toF64Kernel(const DevMesh* __restrict__ meshes,
                                                  const uint4* __restrict__ spans, uint32_t nspans)
{
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    uint32_t i = bl.y * VERT_BLOCK + threadIdx.x;
    if(i >= m.nverts) return;
    float4 a = m.pos32[i];
    m.pos64[i] = make_double4((double)a.x, (double)a.y, (double)a.z, 0.0);
}

__global__ void __launch_bounds__(VERT_BLOCK) resetCountsKernel(const DevMesh* __restrict__ meshes,
                                                  const uint4* __restrict__ spans, uint32_t nspans)
{
    const uint2 bl = blockItem(spans, nspans, blockIdx.x);
    const DevMesh& m = meshes[bl.x];
    uint32_t i = bl.y * VERT_BLOCK + threadIdx.x;
    if(i < m.nverts) m.evalCounts[i] = 0;
}

__global__ void resetSlotsKernel(const DevMesh* __restrict__ meshes, const uint32_t* __restrict__ meshList,
                                 uint32_t nmeshes)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= nmeshes) return;
    const DevMesh& m = meshes[meshList[i]];
    m.passSlot[0] = 0;
    m.passSlot[1] = 0;
    for(int b = 0; b < 20; ++b) m.evalHist[b] = 0;
}

__global__ void __launch_bounds__(VERT_BLOCK) gatherCountsKernel(const int32_t* __restrict__ counts,
                                                                 const uint32_t* __restrict__ slotOf,
                                                                 int32_t* __restrict__ out, uint32_t n)
{
    const uint32_t v = blockIdx.x * VERT_BLOCK + threadIdx.x;
    if(v < n) out[v] = counts[slotOf[v]];
}

hipError_t launchGatherCounts(const int32_t* counts, const uint32_t* slotOf, int32_t* out, uint32_t n,
                              hipStream_t stream)
{
    if(n == 0) return hipSuccess;
    gatherCountsKernel<<<(n + VERT_BLOCK - 1) / VERT_BLOCK, VERT_BLOCK, 0, stream>>>(counts, slotOf, out, n);
    return hipGetLastError();
}

hipError_t launchToF32(const DevMesh* meshes, BlockWork vb, hipStream_t stream)
{
    if(vb.count == 0) return hipSuccess;
    toF32Kernel<<<vb.count, VERT_BLOCK, 0, stream>>>(meshes, vb.spans, vb.nspans);
    return hipGetLastError();
}

hipError_t launchToF64(const DevMesh* meshes, BlockWork vb, hipStream_t stream)
{
    if(vb.count == 0) return hipSuccess;
    toF64Kernel<<<vb.count, VERT_BLOCK, 0, stream>>>(meshes, vb.spans, vb.nspans);
    return hipGetLastError();
}

hipError_t launchResetRun(const DevMesh* meshes, BlockWork vb, const uint32_t* meshList, uint32_t nmeshes,
                          hipStream_t stream)
{
    if(vb.count) resetCountsKernel<<<vb.count, VERT_BLOCK, 0, stream>>>(meshes, vb.spans, vb.nspans);
    if(nmeshes) resetSlotsKernel<<<(nmeshes + 255) / 256, 256, 0, stream>>>(meshes, meshList, nmeshes);
    return hipGetLastError();
}

bool laneFits(const VertexWork& w, int smoother, LaneShape lane, bool fp64, int evaluator, int sampler)
{
    const bool mr = evaluator == gmq::EV_MEAN_RATIO && sampler == gmq::SA_UNIFORM;
    if(fp64) return mr ? laneFits<double, gmq::EV_MEAN_RATIO, gmq::SA_UNIFORM>(w, smoother, lane)
                       : laneFits<double, 1, 1>(w, smoother, lane);
    return mr ? laneFits<float, gmq::EV_MEAN_RATIO, gmq::SA_UNIFORM>(w, smoother, lane)
              : laneFits<float, 1, 1>(w, smoother, lane);
}

__global__ void mathProbeKernel(int op, const double* a, const double* b, double* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    double x = a[i], r;
    switch(op)
    {
    case 0: r = gmq::tsqrt(x); break;
    case 1: r = x / b[i]; break;
    case 2: r = gm_pow23(x); break;
    case 3: r = gm_cbrt_nonneg(x); break;
    case 5: r = gma::powm23(x); break;
    case 6: r = gm_cos(x); break;
    case 7: r = gm_log(x); break;
    case 8: r = gm_pow(x, b[i]); break;
    default: r = gm_exp(x); break;
    }
    out[i] = r;
}

// Issue-rate probe of the FP32 vector instructions the affine kernels spend
// their time in.  Every lane runs THR_ACC independent chains for `iters`
// steps: op 0 the transcendental pair of a^(-2/3) (a = exp2(-|log2 a|), a
// fixed point for a in (0, 1], so the chain stays bounded with no other
// instruction: the negation and the absolute value are operand modifiers),
// op 1 one FMA per step, op 2 both on separate chains (whether the two kinds
// overlap).  Each lane writes the sum of its chains.
constexpr int THR_ACC = 8;

__global__ void __launch_bounds__(256) throughputKernel(int op, int iters, float* out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[THR_ACC], f[THR_ACC];
#pragma unroll
    for(int k = 0; k < THR_ACC; ++k)
    {
        a[k] = 0.25f + 0.0625f * (float)k + 1e-7f * (float)(t & 1023);
        f[k] = a[k];
    }
    const float b = 0.999f, c = 1e-3f;
    for(int i = 0; i < iters; ++i)
    {
        if(op != 1)
#pragma unroll
            for(int k = 0; k < THR_ACC; ++k)
                a[k] = __builtin_amdgcn_exp2f(-__builtin_fabsf(__builtin_amdgcn_logf(a[k])));
        if(op != 0)
#pragma unroll
            for(int k = 0; k < THR_ACC; ++k)
                f[k] = __builtin_fmaf(f[k], b, c);
    }
    float s = 0.0f;
#pragma unroll
    for(int k = 0; k < THR_ACC; ++k) s += a[k] + f[k];
    out[t] = s;
}

hipError_t launchThroughput(int op, int blocks, int iters, float* out, hipStream_t stream)
{
    throughputKernel<<<blocks, 256, 0, stream>>>(op, iters, out);
    return hipGetLastError();
}

int throughputChains() { return THR_ACC; }

hipError_t launchMathProbe(int op, const double* a, const double* b, double* out, int n, hipStream_t stream)
{
    mathProbeKernel<<<(n + 255) / 256, 256, 0, stream>>>(op, a, b, out, n);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    return hipStreamSynchronize(stream);
}

} // namespace gmd
