// FP64 plain form: built without FP contraction so every product and sum rounds as the
// reference's x86-64 build does (bit-exact parity with the oracle).
#include "kernels.inc"

namespace rkdg {
template int stage_persistent_resident<double, false>(int, bool, int);
template hipError_t launch_stage<double, false>(int, bool, bool, const StageArgs<double>&, hipStream_t, int, int);
template hipError_t launch_limiter<double>(int, const LimiterArgs<double>&, hipStream_t, int);
template hipError_t launch_limiter_prepare<double>(int, const LimiterArgs<double>&, hipStream_t, int);
template hipError_t launch_limiter_limit<double>(int, const LimiterArgs<double>&, hipStream_t, int);
template hipError_t launch_limiter_list<double>(int, const LimiterArgs<double>&, const int*, const int*, hipStream_t);
template hipError_t launch_limiter_judge<double>(int, const LimiterArgs<double>&, const JudgeArgs&, hipStream_t, int);
template hipError_t launch_halo<double>(bool, int, int, double*, uint8_t*, const int*, int, double*, hipStream_t);
}  // namespace rkdg

namespace rkdg {

// ESSPRKStepper::operator++ (explicit_ssp_rk_stepper.hpp:46-65) for every ensemble instance, on
// the device so that a captured step graph replays with the right stage times: stage <- (stage+1)
// mod ns, on wrap t += dt and step += 1; stage time t + dt drk[stage].  Built without FP
// contraction, like the host stepper, so stage times match it bit for bit.  The ramp
// tanh(2 (t_s/86400)/ramp_duration) is not evaluated here: instances with a ramp duration read
// it from ramp_tab ([row][stage][instance], row = step - *ramp_base), which the host fills with
// its own tanh, so the ramp equals the single-mesh stepper's bit for bit.
__global__ void batch_advance_kernel(int n, int ns, double* t, const double* dt, const double* rd, int* step, int* stage,
                                     const double* drk, double* tr, unsigned long long* step_ctr,
                                     const double* ramp_tab, const long long* ramp_base, int ramp_rows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int s = stage[i] + 1;
    s = s % ns;
    stage[i] = s;
    if (s == 0) {
        t[i] += dt[i];
        step[i] += 1;
    }
    const double ts = t[i] + dt[i] * drk[s];
    tr[2 * i] = ts;
    const bool no_ramp = fabs(rd[i]) < 2.220446049250313e-16 * 100;
    if (!no_ramp && ramp_tab) {
        const long long row = (long long)step[i] - *ramp_base;
        if (row >= 0 && row < ramp_rows) tr[2 * i + 1] = ramp_tab[((size_t)row * ns + s) * n + i];
    }
    if (i == 0) *step_ctr = (unsigned long long)step[0];
}

__global__ void copy_bytes_kernel(size_t n, const uint8_t* __restrict__ src, uint8_t* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_batch_advance(int n, int ns, double* t, const double* dt, const double* rd, int* step, int* stage,
                                const double* drk, double* tr, unsigned long long* step_ctr, const double* ramp_tab,
                                const long long* ramp_base, int ramp_rows, hipStream_t s) {
    hipLaunchKernelGGL(batch_advance_kernel, dim3((n + 63) / 64), dim3(64), 0, s, n, ns, t, dt, rd, step, stage, drk, tr,
                       step_ctr, ramp_tab, ramp_base, ramp_rows);
    return hipGetLastError();
}

hipError_t launch_copy_bytes(size_t n, const uint8_t* src, uint8_t* dst, hipStream_t s) {
    hipLaunchKernelGGL(copy_bytes_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, src, dst);
    return hipGetLastError();
}

}  // namespace rkdg
