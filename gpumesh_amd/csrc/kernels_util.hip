// kernels_util.hip -- precision conversions and the device math probe.
// Built with -ffp-contract=off like every FP64 kernel.
#include "gm_kernels.h"

namespace gmd
{

__global__ void toF32Kernel(const double4* __restrict__ in, float4* __restrict__ out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    double4 a = in[i];
    out[i] = make_float4((float)a.x, (float)a.y, (float)a.z, 0.0f);
}

__global__ void toF64Kernel(const float4* __restrict__ in, double4* __restrict__ out, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= n) return;
    float4 a = in[i];
    out[i] = make_double4((double)a.x, (double)a.y, (double)a.z, 0.0);
}

hipError_t launchToF32(const DevMesh& m, hipStream_t stream)
{
    if(m.nverts == 0) return hipSuccess;
    toF32Kernel<<<(m.nverts + 255) / 256, 256, 0, stream>>>(m.pos64, m.pos32, m.nverts);
    return hipGetLastError();
}

hipError_t launchToF64(const DevMesh& m, hipStream_t stream)
{
    if(m.nverts == 0) return hipSuccess;
    toF64Kernel<<<(m.nverts + 255) / 256, 256, 0, stream>>>(m.pos32, m.pos64, m.nverts);
    return hipGetLastError();
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
    default: r = gm_exp(x); break;
    }
    out[i] = r;
}

hipError_t launchMathProbe(int op, const double* a, const double* b, double* out, int n, hipStream_t stream)
{
    mathProbeKernel<<<(n + 255) / 256, 256, 0, stream>>>(op, a, b, out, n);
    hipError_t e = hipGetLastError();
    if(e != hipSuccess) return e;
    return hipStreamSynchronize(stream);
}

} // namespace gmd
