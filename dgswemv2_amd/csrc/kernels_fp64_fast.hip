// FP64 fast form: the fast stage arithmetic in FP64, built with FP contraction (fused
// multiply-adds).  Checked against the plain form at 1e-12 relative per-step L2.
#include "kernels.inc"

namespace rkdg {
template int stage_persistent_resident<double, true>(int, bool, int);
template hipError_t launch_stage<double, true>(int, bool, bool, const StageArgs<double>&, hipStream_t, int, int);
}  // namespace rkdg
