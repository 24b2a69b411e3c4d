// Host-callable launchers of the stage and limiter kernels (kernels.inc).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "limiter.hpp"

namespace rkdg {

// FAST selects the fast form; instantiated for <double, false> (kernels_fp64.hip),
// <double, true> (kernels_fp64_fast.hip) and <float, true> (kernels_fp32.hip).
// persistent_grid > 0 (a multiple of 8): the persistent tile-loop kernel on that many work-groups.
template <typename Real, bool FAST>
hipError_t launch_stage(int p, bool wd, bool cf, const StageArgs<Real>& A, hipStream_t s, int block,
                        int persistent_grid = 0);

template <typename Real>
hipError_t launch_limiter(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);

// Ensemble helpers (kernels_fp64.hip): per-instance stepper advance on the device, byte copy.
hipError_t launch_batch_advance(int n, int ns, double* t, const double* dt, const double* rd, int* step, int* stage,
                                const double* drk, double* tr, unsigned long long* step_ctr, hipStream_t s);
hipError_t launch_copy_bytes(size_t n, const uint8_t* src, uint8_t* dst, hipStream_t s);

}  // namespace rkdg
