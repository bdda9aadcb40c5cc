// One storage variant of the per-precision kernels (lb_kernels.h): built
// once per LB_INST_P (LB_FP64, LB_FP32_SHIFTED, LB_FP16_STORAGE,
// LB_FP32_STORAGE) by jyu-lb_amd/build.py and linked into libjyulb_hip.so.
#include "lb_kernels.h"

#ifndef LB_INST_P
#error "LB_INST_P must name the storage variant to instantiate"
#endif

namespace lb {

template void launch_step_rlx<LB_INST_P>(const StepParams&, int, bool, hipStream_t);
template void launch_init<LB_INST_P>(const StepParams&, hipStream_t);
template void launch_post_f<LB_INST_P>(const StepParams&, int, double*, hipStream_t);
#if LB_INST_P == LB_FP64 || LB_INST_P == LB_FP32_SHIFTED
// the adjoint and the dissipation objective run in FP64 or FP32-shifted precision
template void launch_adj_rlx<LB_INST_P>(const StepParams&, const AdjParams&, int, bool, hipStream_t);
template void launch_diss_rlx<LB_INST_P>(const StepParams&, int, bool, hipStream_t);
#endif

}  // namespace lb
