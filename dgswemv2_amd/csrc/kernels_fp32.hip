// FP32 bulk form: built with FP contraction (fused multiply-adds); its error against the FP64
// oracle is measured, not bit-matched.
#include "kernels.inc"

namespace rkdg {
template int stage_persistent_resident<float, true>(int, bool, int);
template hipError_t launch_stage<float, true>(int, bool, bool, const StageArgs<float>&, hipStream_t, int, int);
template hipError_t launch_limiter<float>(int, const LimiterArgs<float>&, hipStream_t, int);
template hipError_t launch_limiter_prepare<float>(int, const LimiterArgs<float>&, hipStream_t, int);
template hipError_t launch_limiter_limit<float>(int, const LimiterArgs<float>&, hipStream_t, int);
template hipError_t launch_limiter_list<float>(int, const LimiterArgs<float>&, const int*, const int*, hipStream_t);
template hipError_t launch_limiter_judge<float>(int, const LimiterArgs<float>&, const JudgeArgs&, hipStream_t, int);
template hipError_t launch_halo<float>(bool, int, int, float*, uint8_t*, const int*, int, float*, hipStream_t);
}  // namespace rkdg
