// dispatch.hip -- run-time selection of the (evaluator, sampler) kernel set.
#include "gm_device.h"

namespace gmd
{

#define GMD_SELECT(EVV, SAV, T, CALL)                                               \
    switch((EVV) * 3 + (SAV))                                                       \
    {                                                                               \
    case 0: return Launchers<T, 0, 0>::CALL;                                        \
    case 1: return Launchers<T, 0, 1>::CALL;                                        \
    case 2: return Launchers<T, 0, 2>::CALL;                                        \
    case 3: return Launchers<T, 1, 0>::CALL;                                        \
    case 4: return Launchers<T, 1, 1>::CALL;                                        \
    case 5: return Launchers<T, 1, 2>::CALL;                                        \
    default: return hipErrorInvalidValue;                                           \
    }

template<typename T>
hipError_t launchSmoothGroup(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler, int smoother,
                             int shape, const uint32_t* nodes, uint32_t count, double moveCoeff,
                             int32_t* evalCounts, int maxPatch, hipStream_t stream)
{
    if(evaluator < 0 || evaluator > 1 || sampler < 0 || sampler > 2) return hipErrorInvalidValue;
    GMD_SELECT(evaluator, sampler, T, smooth(m, c, smoother, shape, nodes, count, moveCoeff, evalCounts, maxPatch, stream))
}

template<typename T>
hipError_t launchGetmeElements(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                               void* tetOut, void* hexOut, hipStream_t stream)
{
    if(evaluator < 0 || evaluator > 1 || sampler < 0 || sampler > 2) return hipErrorInvalidValue;
    GMD_SELECT(evaluator, sampler, T, getmeElements(m, c, tetOut, hexOut, stream))
}

template<typename T>
hipError_t launchGetmeVertices(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                               const uint32_t* nodes, uint32_t count, const void* tetOut, const void* hexOut,
                               int32_t* evalCounts, hipStream_t stream)
{
    if(evaluator < 0 || evaluator > 1 || sampler < 0 || sampler > 2) return hipErrorInvalidValue;
    GMD_SELECT(evaluator, sampler, T, getmeVertices(m, c, nodes, count, tetOut, hexOut, evalCounts, stream))
}

#define GMD_SELECT_EVAL(EVV, SAV, CALL)                                             \
    switch((EVV) * 3 + (SAV))                                                       \
    {                                                                               \
    case 0: return Evaluators<0, 0>::CALL;                                          \
    case 1: return Evaluators<0, 1>::CALL;                                          \
    case 2: return Evaluators<0, 2>::CALL;                                          \
    case 3: return Evaluators<1, 0>::CALL;                                          \
    case 4: return Evaluators<1, 1>::CALL;                                          \
    case 5: return Evaluators<1, 2>::CALL;                                          \
    default: return hipErrorInvalidValue;                                           \
    }

template<>
hipError_t launchEvaluate<double>(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                                  const EvalScratch& s, int pass, hipStream_t stream)
{
    GMD_SELECT_EVAL(evaluator, sampler, evaluateF64(m, c, s, pass, stream))
}

template<>
hipError_t launchEvaluate<float>(const DevMesh& m, const gmq::Crew& c, int evaluator, int sampler,
                                 const EvalScratch& s, int pass, hipStream_t stream)
{
    GMD_SELECT_EVAL(evaluator, sampler, evaluateF32(m, c, s, pass, stream))
}

template hipError_t launchSmoothGroup<double>(const DevMesh&, const gmq::Crew&, int, int, int, int, const uint32_t*,
                                              uint32_t, double, int32_t*, int, hipStream_t);
template hipError_t launchSmoothGroup<float>(const DevMesh&, const gmq::Crew&, int, int, int, int, const uint32_t*,
                                             uint32_t, double, int32_t*, int, hipStream_t);
template hipError_t launchGetmeElements<double>(const DevMesh&, const gmq::Crew&, int, int, void*, void*, hipStream_t);
template hipError_t launchGetmeElements<float>(const DevMesh&, const gmq::Crew&, int, int, void*, void*, hipStream_t);
template hipError_t launchGetmeVertices<double>(const DevMesh&, const gmq::Crew&, int, int, const uint32_t*, uint32_t,
                                                const void*, const void*, int32_t*, hipStream_t);
template hipError_t launchGetmeVertices<float>(const DevMesh&, const gmq::Crew&, int, int, const uint32_t*, uint32_t,
                                               const void*, const void*, int32_t*, hipStream_t);

} // namespace gmd
