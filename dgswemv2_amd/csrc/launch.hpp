// Host-callable launchers of the stage and limiter kernels (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "limiter.hpp"

namespace rkdg {

template <typename Real>
hipError_t launch_stage(int p, bool wd, bool cf, const StageArgs<Real>& A, hipStream_t s, int block);

template <typename Real>
hipError_t launch_limiter(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);

}  // namespace rkdg
