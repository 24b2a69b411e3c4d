// HIP kernels of the RKDG stage engine (gfx950).
//
// stage_kernel: one work-item per element, the fused stage of engine.hpp.  Work-groups are
// remapped so that each of the 8 XCDs walks a contiguous element range: consecutive
// hardware work-group IDs are dispatched round-robin over the XCDs, and without the remap
// the neighbour rows an element gathers (its +-1 and +-2m neighbours on a structured mesh)
// would sit in another XCD's L2.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "limiter.hpp"
#include "launch.hpp"

namespace rkdg {

__device__ __forceinline__ int xcd_element(int per_xcd, int nblk, int& e) {
    const int b = blockIdx.x;
    const int logical = (b % 8) * per_xcd + b / 8;
    e = logical * blockDim.x + threadIdx.x;
    return logical < nblk;
}

template <typename Real, int P, bool WD>
__global__ void __launch_bounds__(256) stage_kernel(StageArgs<Real> A, const Tables<Real, P>* __restrict__ T,
                                                    int per_xcd, int nblk) {
    int e;
    if (!xcd_element(per_xcd, nblk, e)) return;
    if (e >= A.ne) return;
    stage_element<Real, P, WD>(A, *T, e);
}

template <typename Real, int P>
__global__ void __launch_bounds__(256) sl_prepare_kernel(LimiterArgs<Real> L, const Tables<Real, P>* __restrict__ T,
                                                         int per_xcd, int nblk) {
    int e;
    if (!xcd_element(per_xcd, nblk, e)) return;
    if (e >= L.ne) return;
    sl_prepare_element<Real, P>(L, *T, e);
}

__global__ void __launch_bounds__(256) sl_dilate_kernel(int ne, const int* __restrict__ neighbor,
                                                        const uint8_t* __restrict__ in_set, uint8_t* __restrict__ out_set,
                                                        uint8_t* __restrict__ troubled) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    sl_dilate_element(ne, neighbor, in_set, out_set, troubled, e);
}

template <typename Real, int P>
__global__ void __launch_bounds__(256) sl_limit_kernel(LimiterArgs<Real> L, const Tables<Real, P>* __restrict__ T,
                                                       int per_xcd, int nblk) {
    int e;
    if (!xcd_element(per_xcd, nblk, e)) return;
    if (e >= L.ne) return;
    sl_limit_element<Real, P>(L, *T, e);
}

static inline void grid_of(int ne, int block, int& nblk, int& per_xcd, int& launched) {
    nblk = (ne + block - 1) / block;
    per_xcd = (nblk + 7) / 8;
    launched = per_xcd * 8;
}

template <typename Real>
hipError_t launch_stage(int p, bool wd, const StageArgs<Real>& A, const void* tables, hipStream_t s, int block) {
    int nblk, per_xcd, launched;
    grid_of(A.ne, block, nblk, per_xcd, launched);
    if (p == 1) {
        if (wd)
            hipLaunchKernelGGL((stage_kernel<Real, 1, true>), dim3(launched), dim3(block), 0, s, A,
                               (const Tables<Real, 1>*)tables, per_xcd, nblk);
        else
            hipLaunchKernelGGL((stage_kernel<Real, 1, false>), dim3(launched), dim3(block), 0, s, A,
                               (const Tables<Real, 1>*)tables, per_xcd, nblk);
    } else if (p == 2) {
        if (wd)
            hipLaunchKernelGGL((stage_kernel<Real, 2, true>), dim3(launched), dim3(block), 0, s, A,
                               (const Tables<Real, 2>*)tables, per_xcd, nblk);
        else
            hipLaunchKernelGGL((stage_kernel<Real, 2, false>), dim3(launched), dim3(block), 0, s, A,
                               (const Tables<Real, 2>*)tables, per_xcd, nblk);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename Real>
hipError_t launch_limiter(int p, const LimiterArgs<Real>& L, const void* tables, hipStream_t s, int block) {
    int nblk, per_xcd, launched;
    grid_of(L.ne, block, nblk, per_xcd, launched);
    if (p == 1)
        hipLaunchKernelGGL((sl_prepare_kernel<Real, 1>), dim3(launched), dim3(block), 0, s, L,
                           (const Tables<Real, 1>*)tables, per_xcd, nblk);
    else if (p == 2)
        hipLaunchKernelGGL((sl_prepare_kernel<Real, 2>), dim3(launched), dim3(block), 0, s, L,
                           (const Tables<Real, 2>*)tables, per_xcd, nblk);
    else
        return hipErrorInvalidValue;
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    // six neighbour-dilation passes of the troubled set (swe_trouble_check.hpp:106-115)
    const int nb1 = (L.ne + block - 1) / block;
    uint8_t* cur = L.set_a;
    uint8_t* nxt = L.set_b;
    for (int pass = 0; pass < 6; ++pass) {
        hipLaunchKernelGGL(sl_dilate_kernel, dim3(nb1), dim3(block), 0, s, L.ne, L.neighbor, cur, nxt, L.troubled);
        uint8_t* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (p == 1)
        hipLaunchKernelGGL((sl_limit_kernel<Real, 1>), dim3(launched), dim3(block), 0, s, L,
                           (const Tables<Real, 1>*)tables, per_xcd, nblk);
    else
        hipLaunchKernelGGL((sl_limit_kernel<Real, 2>), dim3(launched), dim3(block), 0, s, L,
                           (const Tables<Real, 2>*)tables, per_xcd, nblk);
    return hipGetLastError();
}

template hipError_t launch_stage<double>(int, bool, const StageArgs<double>&, const void*, hipStream_t, int);
template hipError_t launch_stage<float>(int, bool, const StageArgs<float>&, const void*, hipStream_t, int);
template hipError_t launch_limiter<double>(int, const LimiterArgs<double>&, const void*, hipStream_t, int);
template hipError_t launch_limiter<float>(int, const LimiterArgs<float>&, const void*, hipStream_t, int);

}  // namespace rkdg
