// kernels_inst.hip -- instantiates the kernel set of one (precision,
// evaluator, sampler) combination, selected by -DGM_T=<double|float>
// -DGM_EV=<0|1> -DGM_SA=<0|1|2>; the Makefile builds every combination as a
// separate object so the instantiations compile in parallel.  FP64 objects
// are built with -ffp-contract=off and also carry the FP64 mesh-quality
// reduction of their (evaluator, sampler).
#include "gm_kernels.h"

template struct gmd::Launchers<GM_T, GM_EV, GM_SA>;
#if GM_F64
template struct gmd::Evaluators<GM_EV, GM_SA>;
#endif
