// dispatch.hip -- run-time selection of the (evaluator, sampler) kernel set.
#include "gm_device.h"

namespace gmd
{

#define GMD_SELECT(EVV, SAV, T, CALL)                                               \
    switch((EVV) * 4 + (SAV))                                                       \
    {                                                                               \
    case 0: return Launchers<T, 0, 0>::CALL;                                        \
    case 1: return Launchers<T, 0, 1>::CALL;                                        \
    case 2: return Launchers<T, 0, 2>::CALL;                                        \
    case 3: return Launchers<T, 0, 3>::CALL;                                        \
    case 4: return Launchers<T, 1, 0>::CALL;                                        \
    case 5: return Launchers<T, 1, 1>::CALL;                                        \
    case 6: return Launchers<T, 1, 2>::CALL;                                        \
    case 7: return Launchers<T, 1, 3>::CALL;                                        \
    default: return hipErrorInvalidValue;                                           \
    }

#define GMD_SELECT_EVAL(EVV, SAV, CALL)                                             \
    switch((EVV) * 4 + (SAV))                                                       \
    {                                                                               \
    case 0: return Evaluators<0, 0>::CALL;                                          \
    case 1: return Evaluators<0, 1>::CALL;                                          \
    case 2: return Evaluators<0, 2>::CALL;                                          \
    case 3: return Evaluators<0, 3>::CALL;                                          \
    case 4: return Evaluators<1, 0>::CALL;                                          \
    case 5: return Evaluators<1, 1>::CALL;                                          \
    case 6: return Evaluators<1, 2>::CALL;                                          \
    case 7: return Evaluators<1, 3>::CALL;                                          \
    default: return hipErrorInvalidValue;                                           \
    }

template<typename T>
hipError_t launchSmooth(const VertexWork& w, const gmq::Crew& c, int evaluator, int sampler, int smoother,
                        int shape, LaneShape lane, double moveCoeff, hipStream_t stream)
{
    GMD_SELECT(evaluator, sampler, T, smooth(w, c, smoother, shape, lane, moveCoeff, stream))
}

template<typename T>
hipError_t launchGetmeElements(const DevMesh* meshes, BlockWork tets, BlockWork hexs, BlockWork pris,
                               const gmq::Crew& c, int evaluator, int sampler, bool fast, hipStream_t stream)
{
    GMD_SELECT(evaluator, sampler, T, getmeElements(meshes, tets, hexs, pris, c, fast, stream))
}

template<typename T>
hipError_t launchGetmeVertices(const VertexWork& w, const gmq::Crew& c, int evaluator, int sampler, int shape,
                               LaneShape lane, hipStream_t stream)
{
    GMD_SELECT(evaluator, sampler, T, getmeVertices(w, c, shape, lane, stream))
}

template<>
hipError_t launchEvaluate<double>(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                  uint32_t nmeshes, const gmq::Crew& c, int evaluator, int sampler, bool hexes,
                                  hipStream_t stream)
{
    GMD_SELECT_EVAL(evaluator, sampler, evaluateF64(meshes, blocks, meshList, nmeshes, c, hexes, stream))
}

template<>
hipError_t launchEvaluate<float>(const DevMesh* meshes, BlockWork blocks, const uint32_t* meshList,
                                 uint32_t nmeshes, const gmq::Crew& c, int evaluator, int sampler, bool hexes,
                                 hipStream_t stream)
{
    GMD_SELECT_EVAL(evaluator, sampler, evaluateF32(meshes, blocks, meshList, nmeshes, c, hexes, stream))
}

template hipError_t launchSmooth<double>(const VertexWork&, const gmq::Crew&, int, int, int, int, LaneShape, double,
                                         hipStream_t);
template hipError_t launchSmooth<float>(const VertexWork&, const gmq::Crew&, int, int, int, int, LaneShape, double,
                                        hipStream_t);
template hipError_t launchGetmeElements<double>(const DevMesh*, BlockWork, BlockWork, BlockWork, const gmq::Crew&,
                                                int, int, bool, hipStream_t);
template hipError_t launchGetmeElements<float>(const DevMesh*, BlockWork, BlockWork, BlockWork, const gmq::Crew&,
                                               int, int, bool, hipStream_t);
template hipError_t launchGetmeVertices<double>(const VertexWork&, const gmq::Crew&, int, int, int, LaneShape,
                                                hipStream_t);
template hipError_t launchGetmeVertices<float>(const VertexWork&, const gmq::Crew&, int, int, int, LaneShape,
                                               hipStream_t);

} // namespace gmd
