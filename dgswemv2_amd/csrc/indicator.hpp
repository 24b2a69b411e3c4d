// Learned troubled-cell indicator: a 36 -> 64 -> 64 -> 64 -> 1 ReLU MLP that predicts, per
// element, whether the Cockburn-Shu limiter (slope_limiting_kernel, swe_CS_slope_limiter.hpp:88-189)
// will write back a change that matters, so the limiter runs only where the network says so.
// It has no counterpart in the reference; what it predicts is the reference limiter's decision.
//
// Features of element e (36).  Entities k = 0..3: e, then the neighbours across local faces 0..2
// (e itself across a boundary face).  For variable v (ze, qx, qy) and modes i = 0..2:
//   c[v][k][i] = q[v][i][E_k] - (i == 0) q[v][0][e],   m_v = max |c[v][.][.]|
//   x[9k + 3v + i] = c[v][k][i] / m_v                                          -- shape
//   x[3v] = clamp(1 + log10(m_v / s_v) / 8, -1, 1)                             -- size (k = i = 0,
//           an entry that is zero by construction), s_ze = h_s, s_q = h_s sqrt(g h_s),
//           h_s = max(q[ze][0][e] + b_bc[e], 0.01);
//   a variable with m_v <= 1e-16 s_v (no variation above rounding) has shape 0 and size -1.
//
// MFMA layout (v_mfma_f32_16x16x32_bf16, one wave64 per 16 elements).  Lane l = 16 g + n serves
// element n of the tile and entity / K-group g.  A product D[16 x 16] = A[16 x 32] B[32 x 16] takes
// A[r][8g + j] from lane (g, r) and B[8g + j][n] from lane (g, n), j = 0..7, and leaves
// D[4g + jj][n] in lane (g, n), jj = 0..3.  The K index inside a product is a dummy, so each layer
// fixes its own K order ("slots"):
//   layer 1: slot (ks, g, j) = feature 9 g + 8 ks + j if 8 ks + j < 9, else zero: lane (g, n)
//            feeds the nine features of entity g of element n, computed in that lane;
//   layers 2, 3: slot (ks, g, j) = neuron 32 ks + 4 g + j (j < 4) or 32 ks + 16 + 4 g + j - 4:
//            exactly the outputs D of M-tiles 2 ks, 2 ks + 1 that lane (g, n) already holds,
//            so a layer's output feeds the next product with no data movement.
// The weights are packed once in that order: [mt 4][ks 2][lane 64][j 8] bf16 per matrix.
#pragma once

#include <cstdint>

#ifndef RKDG_HD
#define RKDG_HD __host__ __device__
#endif

namespace rkdg {
namespace mlp {

constexpr int NIN = 36, NH = 64;
// flat parameter vector, the order of rkdg_mlp_weights: w1[64][36] b1 w2[64][64] b2 w3[64][64] b3 w4 b4
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH,
              NPARAM = OFF_B4 + 1;
constexpr int PACK = 4 * 2 * 64 * 8;  // bf16 entries of one packed A-operand matrix
// packed matrices: 0 W1, 1 W2, 2 W3 (forward), 3 W2^T, 4 W3^T (back-propagation)
constexpr int NPACKED = 5;
// small per-layer vectors the kernels read from LDS: b1 b2 b3 w4 b4 threshold
constexpr int NVEC = 4 * NH + 2;
// Samples per weight-gradient slice (one work-group each, partial sums combined in slice order):
// 64 for training batches of up to 4096 samples (more work-groups in flight), 256 beyond (bounds
// the partial-gradient buffer of large gradient evaluations).
constexpr int wgrad_slice(int bpad) { return bpad <= 4096 ? 64 : 256; }
constexpr int wgrad_slices(int bpad) { return (bpad + wgrad_slice(bpad) - 1) / wgrad_slice(bpad); }
// Slices any batch of at most bpad samples may need (the partial-gradient buffer's size).
constexpr int wgrad_slices_cap(int bpad) {
    const int a = wgrad_slices(bpad < 4096 ? bpad : 4096), b = wgrad_slices(bpad);
    return a > b ? a : b;
}
constexpr double H_FLOOR = 0.01;
// A write-back "matters" when it changes the state by more than this share of the physical scale:
// 1e-8 for the FP64 forms; 1e-6 for FP32 bulk, the level of that form's own rounding per step.
constexpr double SIG_TOL = 1.0e-8;
constexpr double SIG_TOL_FP32 = 1.0e-6;
constexpr double VAR_FLOOR = 1.0e-16;  // variations at or below this share of the scale are none

RKDG_HD inline int slot_feature(int ks, int g, int j) {
    const int s = 8 * ks + j;
    return s < 9 ? 9 * g + s : -1;
}
RKDG_HD inline int slot_neuron(int ks, int g, int j) { return 32 * ks + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4)); }

// Flat parameter index feeding packed entry i of matrix `which`, or -1 for a zero slot.
RKDG_HD inline int pack_source(int which, int i) {
    const int j = i & 7, lane = (i >> 3) & 63, ks = (i >> 9) & 1, mt = i >> 10;
    const int g = lane >> 4, r = 16 * mt + (lane & 15);
    switch (which) {
        case 0: {
            const int f = slot_feature(ks, g, j);
            return f < 0 ? -1 : OFF_W1 + r * NIN + f;
        }
        case 1: return OFF_W2 + r * NH + slot_neuron(ks, g, j);
        case 2: return OFF_W3 + r * NH + slot_neuron(ks, g, j);
        case 3: return OFF_W2 + slot_neuron(ks, g, j) * NH + r;  // (W2^T)[r][k] = W2[k][r]
        case 4: return OFF_W3 + slot_neuron(ks, g, j) * NH + r;
        default: return -1;
    }
}

// float -> bfloat16 bits, round to nearest even (finite inputs)
RKDG_HD inline uint16_t bf16_bits(float f) {
    union {
        float f;
        uint32_t u;
    } c{f};
    const uint32_t r = c.u + 0x7FFFu + ((c.u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

}  // namespace mlp
}  // namespace rkdg

namespace rkdg {

// Inference / feature pass over one state (Real = the state's storage type).
template <typename Real>
struct GateArgs {
    int ne;
    const Real* q;            // [3][ND][ne] state after the stage and wetting-drying, before limiting
    const uint8_t* wet;       // [ne]
    const int* neighbor;      // [3][ne] by local face, -1 on a boundary
    const double* bath_bc;    // [ne] bathymetry at the barycentre
    double g;
    const uint16_t* wpack;    // packed W1, W2, W3 (bf16 bits), MLP only
    const float* vec;         // b1 b2 b3 w4 b4 threshold (mlp::NVEC floats), MLP only
    double* baryctr;          // [3][ne] barycentre values for the limit pass, or null
    uint8_t* flag;            // [ne] gate (element may be limited), MLP only
    float* logit;             // [ne] or null
    float* feat;              // [36][ne] or null
    uint8_t* ok;              // [ne] element and neighbours wet, or null
    uint8_t* limited;         // [3][ne] cleared (the limit pass sets it where it writes), MLP only
    int* list;                // elements whose gate is open: wave w's in list[w cap ..], MLP only
    int* list_count;          // [waves] how many each wave found, MLP only
    int list_cap;             // gate_grid(ne).cap
};

// Launch shape of the gate (and of the limit pass over its list): work-groups of 4 waves, each
// wave walking 64-element tiles of its XCD's contiguous range; `cap` bounds the tiles (x 64
// elements) one wave can meet, so wave w's list segment is [w cap, (w + 1) cap).
struct GateGrid {
    int ntiles, per_xcd, waves, cap;
};
inline GateGrid gate_grid(int ne) {
    GateGrid g;
    g.ntiles = (ne + 63) / 64;
    g.per_xcd = (g.ntiles + 31) / 32;          // work-groups per XCD: at most two per CU (32 CUs)
    g.per_xcd = g.per_xcd < 64 ? g.per_xcd : 64;
    g.per_xcd = g.per_xcd > 0 ? g.per_xcd : 1;
    g.waves = 8 * g.per_xcd * 4;
    int most = 0;                              // the largest XCD range, in tiles
    for (int x = 0; x < 8; ++x) {
        const int t0 = (int)((long long)g.ntiles * x / 8), t1 = (int)((long long)g.ntiles * (x + 1) / 8);
        most = t1 - t0 > most ? t1 - t0 : most;
    }
    const int stride = g.per_xcd * 4;          // waves per XCD
    g.cap = 64 * ((most + stride - 1) / stride);
    return g;
}

// Dry-run limiter over all elements: labels and the judge's counters.
struct JudgeArgs {
    const uint8_t* flag;       // gate, or null (plain limiter: label capture only)
    uint8_t* label;            // [ne] significant write-back, or null
    float* change;             // [ne] relative change of the write-back (0 where none), or null
    unsigned long long* cnt;   // [8] counters (see rkdg_indicator_stats), or null
    double h_floor, tol;
};

// One optimisation step's buffers of the trainer (batch of B samples, padded to Bpad, a
// multiple of 32; every [rows][Bpad] array is row-major, sample fastest).
struct TrainArgs {
    int B, Bpad;
    const float* X;           // [36][ldx] features
    long long ldx;
    const float* y;           // [ldx] labels 0 / 1
    const int* idx;           // [B] sample indices of this batch
    float pos_weight;
    const uint16_t* wpack;    // the five packed matrices
    const float* theta;       // [NPARAM] master parameters
    uint16_t *xb, *h1, *h2, *dz1, *dz2, *dz3;  // bf16 bits: [48][Bpad], [64][Bpad] x 5
    float* h3;                // [Bpad][64] (sample-major: the w4 gradient walks samples)
    float* dz;                // [Bpad] dL/dlogit
    float* loss;              // [Bpad] weighted per-sample loss
    float* logit;             // [Bpad] or null
};

}  // namespace rkdg
