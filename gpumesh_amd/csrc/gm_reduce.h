// gm_reduce.h -- device helpers of the FP64 mesh-quality reduction shared by
// the evaluation kernels (gm_kernels.h) and the allocation-free evaluation
// (noalloc.hip).
#ifndef GM_REDUCE_H
#define GM_REDUCE_H

#include <hip/hip_runtime.h>

namespace gmd
{

// QualityHistogram's bucket (DataStructures/QualityHistogram.cpp:140-153):
// size_t(value * 20) clamped to [0, 19]; products that truncate to a
// negative integer and NaN land in the last bucket (x86-64 conversion).
__device__ __forceinline__ int histBucket(double value)
{
    double x = value * 20.0;
    if(x != x || x <= -1.0 || x >= 9.2e18) return 19;
    long long t = (long long)x;
    return t < 0 ? 0 : (t > 19 ? 19 : (int)t);
}

// Error-free sum: s + e == a + b exactly (Knuth's TwoSum).
__device__ __forceinline__ void twoSum(double a, double b, double& s, double& e)
{
    s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

} // namespace gmd

#endif // GM_REDUCE_H
