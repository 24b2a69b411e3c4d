// Host-callable launchers of the stage and limiter kernels (kernels.inc).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "indicator.hpp"
#include "limiter.hpp"

namespace rkdg {

// FAST selects the fast form; instantiated for <double, false> (kernels_fp64.hip),
// <double, true> (kernels_fp64_fast.hip) and <float, true> (kernels_fp32.hip).
// persistent_grid > 0 (a multiple of 8): the persistent tile-loop kernel on that many work-groups.
template <typename Real, bool FAST>
hipError_t launch_stage(int p, bool wd, bool cf, const StageArgs<Real>& A, hipStream_t s, int block,
                        int persistent_grid = 0);
// Work-groups of the ensemble's persistent stage kernel (order p, wetting-drying wd) resident on
// one CU at once at this block size; 0 if the query fails.
template <typename Real, bool FAST>
int stage_persistent_resident(int p, bool wd, int block);

template <typename Real>
hipError_t launch_limiter(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);
// The limiter in two parts: barycentres + trouble check + its six dilation passes, then the
// limit pass; and the learned form's limit pass over the gate's list of open elements.
template <typename Real>
hipError_t launch_limiter_prepare(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);
template <typename Real>
hipError_t launch_limiter_limit(int p, const LimiterArgs<Real>& L, hipStream_t s, int block);
template <typename Real>
hipError_t launch_limiter_list(int p, const LimiterArgs<Real>& L, const int* list, const int* count, hipStream_t s);
// Dry run of the limit pass over every element: labels and the judge's counters.
template <typename Real>
hipError_t launch_limiter_judge(int p, const LimiterArgs<Real>& L, const JudgeArgs& J, hipStream_t s, int block);

// Halo exchange of a partitioned mesh: pack (state -> buffer) or unpack (buffer -> state) of n
// listed elements; buffer [nrow + 1][n] (modes, then the wet flag).
template <typename Real>
hipError_t launch_halo(bool pack, int ne, int nrow, Real* q, uint8_t* wet, const int* idx, int n, Real* buf,
                       hipStream_t s);

// Learned indicator (indicator.hip): features (+ inference with mlp_on) over one state.
template <typename Real>
hipError_t launch_gate(int p, const GateArgs<Real>& G, bool mlp_on, hipStream_t s);
// Trainer steps (indicator.hip).
hipError_t launch_mlp_fwdbwd(const TrainArgs& T, bool backward, hipStream_t s);
hipError_t launch_mlp_wgrad(const TrainArgs& T, float* part, float* loss_part, hipStream_t s);
hipError_t launch_mlp_update(int nslices, const float* part, const float* loss_part, float* grad, float* theta,
                             float* m1, float* m2, float lr, float b1, float b2, float eps, float c1, float c2,
                             bool update, float* loss_out, hipStream_t s);
hipError_t launch_mlp_pack(const float* theta, uint16_t* wpack, hipStream_t s);
// Logits of x [36][n] (host) under the flat parameters theta (host), staged through the library's
// fixed device array (no allocation); n <= RKDG_MLP_FORWARD_NOALLOC_MAX.
hipError_t mlp_forward_noalloc(const float* theta, const float* x, int n, float* logit);

// Ensemble helpers (kernels_fp64.hip): per-instance stepper advance on the device, byte copy.
hipError_t launch_batch_advance(int n, int ns, double* t, const double* dt, const double* rd, int* step, int* stage,
                                const double* drk, double* tr, unsigned long long* step_ctr, const double* ramp_tab,
                                const long long* ramp_base, int ramp_rows, hipStream_t s);
hipError_t launch_copy_bytes(size_t n, const uint8_t* src, uint8_t* dst, hipStream_t s);

}  // namespace rkdg
