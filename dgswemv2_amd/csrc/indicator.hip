// Learned troubled-cell indicator kernels (gfx950): the gate (features + BF16-MFMA inference),
// and the trainer's forward / back-propagation, weight-gradient, Adam and packing kernels.
// Layouts and slot orders are described in indicator.hpp.
//
// Built without FP contraction: the gate writes the barycentres the limit pass reads, and they
// must round exactly as the plain limiter's own (sl_prepare_element) do.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/rkdg.h"

#include "engine.hpp"
#include "indicator.hpp"
#include "launch.hpp"

namespace rkdg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float bf16_to_float(__bf16 b) { return (float)b; }

// max(x, 0) in one instruction (v_med3_f32); fmaxf would also canonicalise NaNs
__device__ __forceinline__ float relu(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 3.40282347e38f); }

// One hidden layer of 16 elements: D = W B (+ bias), W packed [mt][ks] (A operands), B in two
// K-steps.  acc[mt][jj] = neuron 16 mt + 4 g + jj of element n; the bias of those four neurons
// is one 16-byte read.
template <bool BIAS>
__device__ __forceinline__ void layer(const bf16x8 (&W)[4][2], const bf16x8& b0, const bf16x8& b1, const float* bias,
                                      int g, f32x4 (&acc)[4]) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        if constexpr (BIAS) c = *reinterpret_cast<const f32x4*>(bias + 16 * mt + 4 * g);
        c = mfma(W[mt][0], b0, c);
        acc[mt] = mfma(W[mt][1], b1, c);
    }
}

// relu and repack D of M-tiles (0, 1) and (2, 3) as the B operands of the next layer's K-steps
__device__ __forceinline__ void relu_pack(const f32x4 (&acc)[4], bf16x8& b0, bf16x8& b1) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        b0[jj] = (__bf16)relu(acc[0][jj]);
        b0[4 + jj] = (__bf16)relu(acc[1][jj]);
        b1[jj] = (__bf16)relu(acc[2][jj]);
        b1[4 + jj] = (__bf16)relu(acc[3][jj]);
    }
}

// as relu_pack, for back-propagated values masked by the forward activations (bf16 > 0)
__device__ __forceinline__ void mask_pack(const f32x4 (&acc)[4], const bf16x8& h0, const bf16x8& h1, bf16x8& b0,
                                          bf16x8& b1) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        b0[jj] = (__bf16)(bf16_to_float(h0[jj]) > 0.0f ? acc[0][jj] : 0.0f);
        b0[4 + jj] = (__bf16)(bf16_to_float(h0[4 + jj]) > 0.0f ? acc[1][jj] : 0.0f);
        b1[jj] = (__bf16)(bf16_to_float(h1[jj]) > 0.0f ? acc[2][jj] : 0.0f);
        b1[4 + jj] = (__bf16)(bf16_to_float(h1[4 + jj]) > 0.0f ? acc[3][jj] : 0.0f);
    }
}

__device__ __forceinline__ uint16_t bits(__bf16 b) { return __builtin_bit_cast(uint16_t, b); }

// Orders a wave's LDS writes before its later LDS reads of other lanes' data (no work-group barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// store a packed pair (neurons of M-tiles 2 ks, 2 ks + 1) as rows of a [64][ld] bf16 array
__device__ __forceinline__ void store_rows(uint16_t* dst, long long ld, int col, int g, const bf16x8& b0,
                                           const bf16x8& b1) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        dst[(long long)(0 + 4 * g + jj) * ld + col] = bits(b0[jj]);
        dst[(long long)(16 + 4 * g + jj) * ld + col] = bits(b0[4 + jj]);
        dst[(long long)(32 + 4 * g + jj) * ld + col] = bits(b1[jj]);
        dst[(long long)(48 + 4 * g + jj) * ld + col] = bits(b1[4 + jj]);
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Gate.  A work-group of 4 waves takes 256 consecutive elements per iteration (block-uniform trip
// count; XCD x walks the contiguous range [x T/8, (x+1) T/8) of 256-element tiles so neighbour
// gathers stay in its L2).  Phase 1, one lane per element: the 36 features (36 independent
// gathers per lane), the may-act flag, the barycentres; the features go to LDS as bf16 rows of
// four 16-slot entity blocks.  Phase 2, per 16 elements: lane (g, n) reads entity g's slots of
// element n as its two B operands and the network runs as 24 MFMA products with the weights
// resident in VGPRs; the logit is reduced over the four lanes of each element.  Phase 3, one
// lane per element: the gate flag.
constexpr int kGateRow = 4 * 16 + 8;  // bf16 per element row in LDS (padded: 144 B)

// Phase-1 inputs of one element: the modes 0..2 of the element and its three neighbours, their
// wet flags, the bathymetry at the barycentre.
template <typename Real>
struct GateIn {
    Real val[36];
    uint8_t wet[4];
    double bath;
};

template <typename Real, int P>
__device__ __forceinline__ void gate_load(const GateArgs<Real>& G, int e, GateIn<Real>& in) {
    constexpr int ND = Sizes<P>::ND;
    const int ne = G.ne;
    const int ec = e < ne ? e : ne - 1;
    int E[4];
    E[0] = ec;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int nb = ldc(G.neighbor, ne, k, ec);
        E[k + 1] = nb >= 0 ? nb : ec;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int i = 0; i < 3; ++i) in.val[9 * k + 3 * v + i] = ldc(G.q, ne, v * ND + i, E[k]);
        in.wet[k] = G.wet[E[k]];
    }
    in.bath = ldc(G.bath_bc, ne, 0, ec);
}

// Gate.  Each wave walks 64-element tiles on its own (XCD x's waves take the contiguous range
// [x T/8, (x+1) T/8) so neighbour gathers stay in its L2; a wave only reads LDS rows it wrote, so
// there is no work-group barrier).  Phase 1, one lane per element: the 36 features, the may-act
// flag, the barycentres; the features go to LDS as bf16 rows of four 16-slot entity blocks.
// Phase 2, per 16 elements: lane (g, n) reads entity g's slots of element n as its two B operands
// and the network runs as 24 MFMA products with the weights resident in VGPRs; the logit is
// reduced over the four lanes of each element.  Phase 3, one lane per element: the gate flag and
// the list.  With FP32 state the next tile's inputs are loaded before phase 2, so their latency
// hides behind the products (two dependent gathers per tile otherwise go uncovered).
template <typename Real, int P, bool MLP>
__global__ void __launch_bounds__(256) gate_kernel(GateArgs<Real> G, int ntiles) {
    constexpr bool kPrefetch = MLP && sizeof(Real) == 4;
    __shared__ __attribute__((aligned(16))) float s_vec[mlp::NVEC];
    __shared__ __attribute__((aligned(16))) uint16_t s_x[MLP ? 256 * kGateRow : 1];
    __shared__ float s_z[MLP ? 256 : 1];
    __shared__ uint8_t s_ok[MLP ? 256 : 1];
    if constexpr (MLP) {
        for (int i = threadIdx.x; i < mlp::NVEC; i += blockDim.x) s_vec[i] = G.vec[i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bf16x8 W[3][4][2];
    if constexpr (MLP) {
        const bf16x8* pw = reinterpret_cast<const bf16x8*>(G.wpack);
#pragma unroll
        for (int L = 0; L < 3; ++L)
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) W[L][mt][ks] = pw[((L * 4 + mt) * 2 + ks) * 64 + lane];
    }
    const int wpb = blockDim.x >> 6;
    const int xcd = blockIdx.x & 7, per = (gridDim.x >> 3) * wpb;
    const int t0 = (int)((long long)ntiles * xcd / 8), t1 = (int)((long long)ntiles * (xcd + 1) / 8);
    const int ne = G.ne;
    int tile = t0 + (int)(blockIdx.x >> 3) * wpb + wave;
    const int wid = blockIdx.x * wpb + wave;  // this wave's list segment
    int nopen = 0;
    GateIn<Real> in;
    if (kPrefetch && tile < t1) gate_load<Real, P>(G, tile * 64 + lane, in);
    for (; tile < t1; tile += per) {
        // ---- phase 1: one element per lane
        const int e = tile * 64 + lane;
        const bool valid = e < ne;
        if (!kPrefetch) gate_load<Real, P>(G, e, in);
        float x[36];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int v = 0; v < 3; ++v)
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const Real val = in.val[9 * k + 3 * v + i];
                    x[9 * k + 3 * v + i] = (float)(i == 0 ? val - in.val[3 * v] : val);
                }
        const bool okw = (in.wet[0] & in.wet[1] & in.wet[2] & in.wet[3]) != 0;
        // physical scales; a variation at or below 1e-16 of its scale carries no shape
        float h = (float)((double)in.val[0] + in.bath);
        h = h > (float)mlp::H_FLOOR ? h : (float)mlp::H_FLOOR;
        const float sq = h * sqrtf((float)G.g * h);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            float m = 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int i = 0; i < 3; ++i) m = fmaxf(m, fabsf(x[9 * k + 3 * v + i]));
            const float r = m / (v == 0 ? h : sq);
            const bool varies = r > (float)mlp::VAR_FLOOR;
            const float inv = varies ? 1.0f / m : 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int i = 0; i < 3; ++i) x[9 * k + 3 * v + i] *= inv;
            // log10 r = log2 r * log10(2), from the hardware log2
            x[3 * v] = varies ? fminf(1.0f, fmaxf(-1.0f, 1.0f + __builtin_amdgcn_logf(r) * (0.30102999566f * 0.125f))) : -1.0f;
        }
        if (valid) {
            if (G.feat)
#pragma unroll
                for (int f = 0; f < 36; ++f) stc(G.feat, ne, f, e, x[f]);
            if (G.ok) G.ok[e] = okw ? 1 : 0;
            if (G.baryctr && in.wet[0])  // as sl_prepare_element
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double m[3] = {(double)in.val[3 * k], (double)in.val[3 * k + 1], (double)in.val[3 * k + 2]};
                    double ql[3];
                    to_vertices(m, ql);
                    stc(G.baryctr, ne, k, e, (ql[0] * (1.0 / 3.0) + ql[1] * (1.0 / 3.0)) + ql[2] * (1.0 / 3.0));
                }
        }
        if constexpr (MLP) {
            // features -> LDS row threadIdx.x: entity g at bf16 offset 16 g, slots 0..8, zeros 9..15
            uint16_t* row = s_x + threadIdx.x * kGateRow;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                bf16x8 a, b;
#pragma unroll
                for (int j = 0; j < 8; ++j) a[j] = (__bf16)x[9 * g + j];
                b[0] = (__bf16)x[9 * g + 8];
#pragma unroll
                for (int j = 1; j < 8; ++j) b[j] = (__bf16)0.0f;
                *reinterpret_cast<bf16x8*>(row + 16 * g) = a;
                *reinterpret_cast<bf16x8*>(row + 16 * g + 8) = b;
            }
            s_ok[threadIdx.x] = okw ? 1 : 0;
            wave_sync();
            if (kPrefetch && tile + per < t1) gate_load<Real, P>(G, (tile + per) * 64 + lane, in);
            // ---- phase 2: the wave's 64 elements as four 16-element products
            const int g = lane >> 4, n = lane & 15;
#pragma unroll 1
            for (int nt = 0; nt < 4; ++nt) {
                const int r = wave * 64 + nt * 16 + n;
                const uint16_t* rr = s_x + r * kGateRow + 16 * g;
                bf16x8 b0 = *reinterpret_cast<const bf16x8*>(rr);
                bf16x8 b1 = *reinterpret_cast<const bf16x8*>(rr + 8);
                f32x4 acc[4];
                layer<true>(W[0], b0, b1, s_vec, g, acc);
                relu_pack(acc, b0, b1);
                layer<true>(W[1], b0, b1, s_vec + 64, g, acc);
                relu_pack(acc, b0, b1);
                layer<true>(W[2], b0, b1, s_vec + 128, g, acc);
                float part = 0.0f;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const f32x4 w4 = *reinterpret_cast<const f32x4*>(s_vec + 192 + 16 * mt + 4 * g);
#pragma unroll
                    for (int jj = 0; jj < 4; ++jj) part = fmaf(w4[jj], relu(acc[mt][jj]), part);
                }
                part += __shfl_xor(part, 16);
                part += __shfl_xor(part, 32);
                if (g == 0) s_z[r] = part + s_vec[256];
            }
            wave_sync();
            // ---- phase 3: one element per lane; open elements go to this wave's list segment
            const float z = s_z[threadIdx.x];
            const bool open = valid && s_ok[threadIdx.x] && z > s_vec[257];
            if (valid) {
                if (G.logit) G.logit[e] = z;
                G.flag[e] = open ? 1 : 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) G.limited[(size_t)k * ne + e] = 0;
            }
            const unsigned long long bal = __ballot(open);
            if (open) G.list[(size_t)wid * G.list_cap + nopen + __popcll(bal & ((1ull << lane) - 1ull))] = e;
            nopen += __popcll(bal);
        }
    }
    if (MLP && lane == 0) G.list_count[wid] = nopen;
}

template <typename Real>
hipError_t launch_gate(int p, const GateArgs<Real>& G, bool mlp_on, hipStream_t s) {
    if (G.ne <= 0) return hipSuccess;
    if (mlp_on && (!G.wpack || !G.vec || !G.flag || !G.limited || !G.list || !G.list_count)) return hipErrorInvalidValue;
    const GateGrid gg = gate_grid(G.ne);
    if (mlp_on && G.list_cap != gg.cap) return hipErrorInvalidValue;
    const int ntiles = gg.ntiles, per = gg.per_xcd;
    const dim3 grid(8 * per), block(256);
    if (p == 1) {
        if (mlp_on)
            hipLaunchKernelGGL((gate_kernel<Real, 1, true>), grid, block, 0, s, G, ntiles);
        else
            hipLaunchKernelGGL((gate_kernel<Real, 1, false>), grid, block, 0, s, G, ntiles);
    } else if (p == 2) {
        if (mlp_on)
            hipLaunchKernelGGL((gate_kernel<Real, 2, true>), grid, block, 0, s, G, ntiles);
        else
            hipLaunchKernelGGL((gate_kernel<Real, 2, false>), grid, block, 0, s, G, ntiles);
    } else if (p == 3) {
        if (mlp_on)
            hipLaunchKernelGGL((gate_kernel<Real, 3, true>), grid, block, 0, s, G, ntiles);
        else
            hipLaunchKernelGGL((gate_kernel<Real, 3, false>), grid, block, 0, s, G, ntiles);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_gate<float>(int, const GateArgs<float>&, bool, hipStream_t);
template hipError_t launch_gate<double>(int, const GateArgs<double>&, bool, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Trainer.  Forward + back-propagation of 16 samples per wave (the weights of all five packed
// matrices in LDS, 40 KB per work-group); every activation and back-propagated value that the
// weight gradients need is written as rows of [rows][Bpad] arrays.
template <bool BACKWARD>
__global__ void __launch_bounds__(256) mlp_fwdbwd_kernel(TrainArgs T) {
    __shared__ bf16x8 s_w[mlp::NPACKED * mlp::PACK / 8];
    __shared__ __attribute__((aligned(16))) float s_vec[4 * mlp::NH + 1];
    {
        const bf16x8* pw = reinterpret_cast<const bf16x8*>(T.wpack);
        const int nw = (BACKWARD ? mlp::NPACKED : 3) * mlp::PACK / 8;
        for (int i = threadIdx.x; i < nw; i += blockDim.x) s_w[i] = pw[i];
        for (int i = threadIdx.x; i < 3 * mlp::NH; i += blockDim.x)
            s_vec[i] = T.theta[(i < 64 ? mlp::OFF_B1 : i < 128 ? mlp::OFF_B2 - 64 : mlp::OFF_B3 - 128) + i];
        for (int i = threadIdx.x; i < mlp::NH + 1; i += blockDim.x) s_vec[192 + i] = T.theta[mlp::OFF_W4 + i];
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (tile * 16 >= T.Bpad) return;  // whole wave
    const int p = tile * 16 + n;
    const bool valid = p < T.B;
    const long long sidx = valid ? (T.idx ? (long long)T.idx[p] : (long long)p) : 0;
    float x[9];
#pragma unroll
    for (int s = 0; s < 9; ++s) x[s] = valid ? T.X[(long long)(9 * g + s) * T.ldx + sidx] : 0.0f;
    bf16x8 b0, b1;
#pragma unroll
    for (int j = 0; j < 8; ++j) b0[j] = (__bf16)x[j];
    b1[0] = (__bf16)x[8];
#pragma unroll
    for (int j = 1; j < 8; ++j) b1[j] = (__bf16)0.0f;
    if (BACKWARD) {
#pragma unroll
        for (int s = 0; s < 8; ++s) T.xb[(long long)(9 * g + s) * T.Bpad + p] = bits(b0[s]);
        T.xb[(long long)(9 * g + 8) * T.Bpad + p] = bits(b1[0]);
#pragma unroll
        for (int s = 0; s < 3; ++s) T.xb[(long long)(36 + 3 * g + s) * T.Bpad + p] = 0;
    }
    auto wmat = [&](int which, bf16x8 (&W)[4][2]) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) W[mt][ks] = s_w[which * (mlp::PACK / 8) + (mt * 2 + ks) * 64 + lane];
    };
    bf16x8 W[4][2];
    f32x4 acc[4];
    bf16x8 h1a, h1b, h2a, h2b;
    wmat(0, W);
    layer<true>(W, b0, b1, s_vec, g, acc);
    relu_pack(acc, h1a, h1b);
    wmat(1, W);
    layer<true>(W, h1a, h1b, s_vec + 64, g, acc);
    relu_pack(acc, h2a, h2b);
    wmat(2, W);
    layer<true>(W, h2a, h2b, s_vec + 128, g, acc);
    float h3[4][4];
    float part = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
            h3[mt][jj] = fmaxf(acc[mt][jj], 0.0f);
            part += s_vec[192 + 16 * mt + 4 * g + jj] * h3[mt][jj];
        }
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    const float z = part + s_vec[256];
    if (g == 0 && T.logit) T.logit[p] = valid ? z : 0.0f;
    if (!BACKWARD) return;
    store_rows(T.h1, T.Bpad, p, g, h1a, h1b);
    store_rows(T.h2, T.Bpad, p, g, h2a, h2b);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) T.h3[(long long)p * mlp::NH + 16 * mt + 4 * g + jj] = h3[mt][jj];
    // weighted BCE on the logit, mean over the batch
    const float yv = valid ? T.y[sidx] : 0.0f;
    const float w = valid ? (yv > 0.5f ? T.pos_weight : 1.0f) : 0.0f;
    const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
    const float sig = 1.0f / (1.0f + expf(-z));
    const float dzv = w * (sig - yv) / (float)T.B;
    if (g == 0) {
        T.loss[p] = w * (sp - yv * z);
        T.dz[p] = dzv;
    }
    // dZ3 = w4 dz (h3 > 0)
    f32x4 d3[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) d3[mt][jj] = h3[mt][jj] > 0.0f ? s_vec[192 + 16 * mt + 4 * g + jj] * dzv : 0.0f;
    bf16x8 d0, d1;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
        d0[jj] = (__bf16)d3[0][jj];
        d0[4 + jj] = (__bf16)d3[1][jj];
        d1[jj] = (__bf16)d3[2][jj];
        d1[4 + jj] = (__bf16)d3[3][jj];
    }
    store_rows(T.dz3, T.Bpad, p, g, d0, d1);
    // dZ2 = (W3^T dZ3) (H2 > 0)
    wmat(4, W);
    layer<false>(W, d0, d1, nullptr, g, acc);
    mask_pack(acc, h2a, h2b, d0, d1);
    store_rows(T.dz2, T.Bpad, p, g, d0, d1);
    // dZ1 = (W2^T dZ2) (H1 > 0)
    wmat(3, W);
    layer<false>(W, d0, d1, nullptr, g, acc);
    mask_pack(acc, h1a, h1b, d0, d1);
    store_rows(T.dz1, T.Bpad, p, g, d0, d1);
}

// Weight gradients over one slice of slen samples (work-group = slice): wave 0 dW1 = dZ1 Xb^T
// (64 x 48, columns 36.. are zero padding), wave 1 dW2 = dZ2 H1^T, wave 2 dW3 = dZ3 H2^T, each as
// MFMA products whose K is the sample; bias gradients are the row sums of the bf16 dZ; wave 3
// dW4 = sum dz h3, db4 = sum dz and the slice's loss.  Partial sums go to part[slice][NPARAM]
// and loss_part[slice] and are added in slice order by the update kernel (deterministic).
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(TrainArgs T, float* part, float* loss_part, int slen) {
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15, role = threadIdx.x >> 6;
    const int slice = blockIdx.x, k_begin = slice * slen;
    const int k_end = k_begin + slen < T.Bpad ? k_begin + slen : T.Bpad;
    float* out = part + (size_t)slice * mlp::NPARAM;
    if (role < 3) {
        const uint16_t* A = role == 0 ? T.dz1 : role == 1 ? T.dz2 : T.dz3;
        const uint16_t* Bm = role == 0 ? T.xb : role == 1 ? T.h1 : T.h2;
        const int nnt = role == 0 ? 3 : 4;
        f32x4 acc[4][4];
        float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = k_begin; k0 < k_end; k0 += 32) {
            bf16x8 a[4], b[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                a[mt] = *reinterpret_cast<const bf16x8*>(A + (long long)(16 * mt + n) * T.Bpad + k0 + 8 * g);
#pragma unroll
                for (int j = 0; j < 8; ++j) bsum[mt] += bf16_to_float(a[mt][j]);
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (nt < nnt) b[nt] = *reinterpret_cast<const bf16x8*>(Bm + (long long)(16 * nt + n) * T.Bpad + k0 + 8 * g);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    if (nt < nnt) acc[mt][nt] = mfma(a[mt], b[nt], acc[mt][nt]);
        }
        const int offw = role == 0 ? mlp::OFF_W1 : role == 1 ? mlp::OFF_W2 : mlp::OFF_W3;
        const int offb = role == 0 ? mlp::OFF_B1 : role == 1 ? mlp::OFF_B2 : mlp::OFF_B3;
        const int ncol = role == 0 ? mlp::NIN : mlp::NH;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                if (nt < nnt) {
                    const int col = 16 * nt + n;
                    if (col < ncol)
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) out[offw + (16 * mt + 4 * g + jj) * ncol + col] = acc[mt][nt][jj];
                }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            float s = bsum[mt];
            s += __shfl_xor(s, 16);
            s += __shfl_xor(s, 32);
            if (g == 0) out[offb + 16 * mt + n] = s;
        }
    } else {
        // eight independent partial sums, so eight sample rows are in flight at a time (a single
        // dependent chain waits out one load latency per sample); combined in a fixed order
        float w8[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        for (int k = k_begin; k < k_end; k += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < k_end) w8[u] += T.dz[k + u] * T.h3[(long long)(k + u) * mlp::NH + lane];
        }
        const float w4 = ((w8[0] + w8[1]) + (w8[2] + w8[3])) + ((w8[4] + w8[5]) + (w8[6] + w8[7]));
        out[mlp::OFF_W4 + lane] = w4;
        float db = 0.0f, ls = 0.0f;
        for (int k = k_begin + lane; k < k_end; k += 64) db += T.dz[k], ls += T.loss[k];
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) db += __shfl_xor(db, m), ls += __shfl_xor(ls, m);
        if (lane == 0) out[mlp::OFF_B4] = db, loss_part[slice] = ls;
    }
}

// Sums the slices' partial gradients in order; with `update`, one Adam step on the FP32 master
// parameters (bias corrections c1 = 1/(1 - beta1^t), c2 = 1/(1 - beta2^t) from the host).
__global__ void mlp_update_kernel(int nslices, const float* part, const float* loss_part, float* grad, float* theta,
                                  float* m1, float* m2, float lr, float b1, float b2, float eps, float c1, float c2,
                                  int update, float* loss_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && loss_out) {
        float l = 0.0f;
        for (int s = 0; s < nslices; ++s) l += loss_part[s];
        *loss_out = l;
    }
    if (i >= mlp::NPARAM) return;
    float gs = 0.0f;
    for (int s = 0; s < nslices; ++s) gs += part[(size_t)s * mlp::NPARAM + i];
    if (grad) grad[i] = gs;
    if (!update) return;
    const float m = b1 * m1[i] + (1.0f - b1) * gs;
    const float v = b2 * m2[i] + (1.0f - b2) * gs * gs;
    m1[i] = m, m2[i] = v;
    theta[i] -= lr * (m * c1) / (sqrtf(v * c2) + eps);
}

// bf16 A operands of the five matrices from the master parameters
__global__ void mlp_pack_kernel(const float* theta, uint16_t* wpack) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mlp::NPACKED * mlp::PACK) return;
    const int src = mlp::pack_source(i / mlp::PACK, i % mlp::PACK);
    wpack[i] = src < 0 ? (uint16_t)0 : mlp::bf16_bits(theta[src]);
}

hipError_t launch_mlp_fwdbwd(const TrainArgs& T, bool backward, hipStream_t s) {
    if (T.Bpad % 32 || T.B > T.Bpad || T.B < 0) return hipErrorInvalidValue;
    const int tiles = T.Bpad / 16, blocks = (tiles + 3) / 4;
    if (blocks == 0) return hipSuccess;
    if (backward)
        hipLaunchKernelGGL(mlp_fwdbwd_kernel<true>, dim3(blocks), dim3(256), 0, s, T);
    else
        hipLaunchKernelGGL(mlp_fwdbwd_kernel<false>, dim3(blocks), dim3(256), 0, s, T);
    return hipGetLastError();
}

hipError_t launch_mlp_wgrad(const TrainArgs& T, float* part, float* loss_part, hipStream_t s) {
    if (T.Bpad % 32) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(mlp::wgrad_slices(T.Bpad)), dim3(256), 0, s, T, part, loss_part,
                       mlp::wgrad_slice(T.Bpad));
    return hipGetLastError();
}

hipError_t launch_mlp_update(int nslices, const float* part, const float* loss_part, float* grad, float* theta,
                             float* m1, float* m2, float lr, float b1, float b2, float eps, float c1, float c2,
                             bool update, float* loss_out, hipStream_t s) {
    hipLaunchKernelGGL(mlp_update_kernel, dim3((mlp::NPARAM + 255) / 256), dim3(256), 0, s, nslices, part, loss_part,
                       grad, theta, m1, m2, lr, b1, b2, eps, c1, c2, update ? 1 : 0, loss_out);
    return hipGetLastError();
}

hipError_t launch_mlp_pack(const float* theta, uint16_t* wpack, hipStream_t s) {
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((mlp::NPACKED * mlp::PACK + 255) / 256), dim3(256), 0, s, theta, wpack);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------------------------
// Allocation-free inference (rkdg_mlp_forward_noalloc).  The staging that rkdg_mlp_forward keeps in
// a trainer's device buffers (master parameters, packed bf16 weights, features, logits) lives here
// in one file-scope device array of fixed size.  Each GPU holds its own copy of the array, at its
// own address: the address is looked up once per device with hipGetSymbolAddress and kept, and only
// that address is handed to kernels and copies (in host code the array's name is a stand-in, not a
// device address).  Calls in one process share the array and are serialised by a lock; every
// region a call reads is written by that call first, so nothing needs clearing.
namespace {
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr int kFwdMax = RKDG_MLP_FORWARD_NOALLOC_MAX;
static_assert(kFwdMax % 32 == 0, "the forward kernel writes logits up to the 32-padded sample count");
constexpr size_t kOffTheta = 0;
constexpr size_t kOffWpack = al256(kOffTheta + (size_t)mlp::NPARAM * sizeof(float));
constexpr size_t kOffX = al256(kOffWpack + (size_t)mlp::NPACKED * mlp::PACK * sizeof(uint16_t));
constexpr size_t kOffLogit = al256(kOffX + (size_t)mlp::NIN * kFwdMax * sizeof(float));
constexpr size_t kScratchBytes = al256(kOffLogit + (size_t)kFwdMax * sizeof(float));
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

hipError_t mlp_forward_noalloc(const float* theta_h, const float* x, int n, float* logit) {
    if (n < 1 || n > kFwdMax) return hipErrorInvalidValue;
    static std::mutex mu;
    static void* base[kMaxDevices] = {};
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (!base[dev]) {  // once per device, with that device current
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        base[dev] = ptr;
    }
    unsigned char* s = static_cast<unsigned char*>(base[dev]);
    float* theta = reinterpret_cast<float*>(s + kOffTheta);
    uint16_t* wpack = reinterpret_cast<uint16_t*>(s + kOffWpack);
    float* X = reinterpret_cast<float*>(s + kOffX);
    float* lg = reinterpret_cast<float*>(s + kOffLogit);
    e = hipMemcpy(theta, theta_h, (size_t)mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = launch_mlp_pack(theta, wpack, nullptr);
    if (e == hipSuccess) e = hipMemcpy(X, x, (size_t)mlp::NIN * n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    TrainArgs T;
    memset(&T, 0, sizeof T);
    T.B = n, T.Bpad = (n + 31) / 32 * 32, T.X = X, T.ldx = n, T.idx = nullptr, T.pos_weight = 1.0f;
    T.wpack = wpack, T.theta = theta, T.logit = lg;
    e = launch_mlp_fwdbwd(T, false, nullptr);
    if (e == hipSuccess) e = hipMemcpy(logit, lg, (size_t)n * sizeof(float), hipMemcpyDeviceToHost);
    return e;
}

}  // namespace rkdg
