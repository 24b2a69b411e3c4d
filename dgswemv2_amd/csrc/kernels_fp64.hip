// FP64 plain form: built without FP contraction so every product and sum rounds as the
// reference's x86-64 build does (bit-exact parity with the oracle).
#include "kernels.inc"

namespace rkdg {
template hipError_t launch_stage<double, false>(int, bool, bool, const StageArgs<double>&, hipStream_t, int);
template hipError_t launch_limiter<double>(int, const LimiterArgs<double>&, hipStream_t, int);
}  // namespace rkdg
