// Host-callable launchers of the stage and limiter kernels (kernels.inc).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "limiter.hpp"

namespace rkdg {

// FAST selects the fast form; instantiated for <double, false> (kernels_fp64.hip),
// <double, true> (kernels_fp64_fast.hip) and <float, true> (kernels_fp32.hip).
template <typename Real, bool FAST>
hipError_t launch_stage(int p, bool wd, bool cf, const StageArgs<Real>& A, hipStream_t s, int block);

template <typename Real>
hipError_t launch_limiter(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);

}  // namespace rkdg
