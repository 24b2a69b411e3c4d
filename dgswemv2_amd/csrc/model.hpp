// Host-side model of one mesh: connectivity in the reference's iteration order, FP64
// geometry, reference-element tables, stepper coefficients and the initial state, laid out
// element-fastest for upload to HBM.
#pragma once

#include <string>
#include <vector>

#include "../../include/rkdg.h"
#include "tables.hpp"

namespace rkdg {

// ESSPRKStepper semantics (simulation/stepper/explicit_ssp_rk_stepper.{hpp,cpp}).
struct Stepper {
    int nstages = 0, order = 0;
    double dt = 0;
    int step = 0, stage = 0;
    double t = 0, ramp_duration = 0, ramp = 1;
    std::vector<std::vector<double>> ark, brk;  // [nstages][nstages] (zero beyond the ragged rows)
    std::vector<double> drk;
    bool init(int ns, int ord, double dt_, double ramp_days);
    double stage_time() const { return t + dt * drk[stage]; }
    void advance();
};

struct Model {
    int p = 0, ndof = 0, ngp = 0, ngpe = 0;
    int ne = 0, nedge = 0, nintf = 0;
    Master master;
    rkdg_params prm{};
    double h_o = 0.1, h_thr = 1.0e-6, sl_M = 1.0e-8, sl_nu = 1.5;

    // per element (SoA, element fastest)
    std::vector<double> geom;       // [5][ne] Jinv00 Jinv01 Jinv10 Jinv11 |detJ|
    std::vector<double> bath;       // [ndof][ne]
    std::vector<double> vxy;        // [6][ne] x0 x1 x2 y0 y1 y2 (after the CCW fix)
    std::vector<double> bath_vrtx;  // [3][ne]
    std::vector<double> gmn2;       // [ne]
    bool reference_face_order = true;
    bool closed_forms = false;      // manufactured source or function boundaries present
    std::vector<int> nb;            // [3][ne] per slot: neighbour, or -1-bc
    std::vector<int> edge;          // [3][ne] per slot: edge record
    std::vector<uint32_t> meta;     // [ne] per slot s: own face bits 4s..4s+1, neighbour face 4s+2..4s+3,
                                    // levee flag bit 12+s
    std::vector<int> neighbor;      // [3][ne] by local face: neighbour or -1 (limiter / dilation)
    // limiter geometry (swe_pre_init_data.hpp:252-326), [k][ne]
    std::vector<double> sl_median;  // [6][ne] mx0 my0 mx1 my1 mx2 my2
    std::vector<double> sl_alpha;   // [6][ne] a0_0 a0_1 a1_0 a1_1 a2_0 a2_1
    std::vector<double> sl_rsq;     // [3][ne]
    std::vector<double> sl_len;     // [3][ne] edge lengths (GetLengths)
    std::vector<double> sl_radius;  // [ne]
    std::vector<double> sl_bath_bc; // [ne] bathymetry at the barycentre
    std::vector<int> sl_face_slot;  // [3][ne] local face -> slot

    // per edge record: interfaces (in-side geometry) then boundaries (own geometry)
    std::vector<double> egeom;      // [3][nedge] nx ny sJ
    std::vector<double> sp_vol;     // [ngp][ne] spherical correction sp at the volume points (empty: sp = 1)
    std::vector<double> sp_edge;    // [ngpe][nedge] sp at the edge points, in-side order (empty: sp = 1)
    std::vector<int> eforce;        // [nedge] tide/flow forcing record, levee record, or -1
    std::vector<double> levee_gp;   // [nlevee][3][ngpe] H_barrier, C_subcritical, C_supercritical at in-side points
    int ncon = 0;
    std::vector<double> freq, ffact, eqarg;
    std::vector<double> famp, fphase;  // [nforced][ncon][ngpe]

    // initial state
    std::vector<double> q0;         // [3][ndof][ne]
    std::vector<uint8_t> wet0;      // [ne]

    Stepper stepper;

    // Ensemble layout (set by merge_models): instance i owns elements [base[i], base[i] + ne[i]),
    // padded to a multiple of kEnsembleTile; padding elements are dry, without neighbours.
    struct Ensemble {
        int n = 0;
        std::vector<int> base, ne;              // per instance
        std::vector<int> tile_inst;             // per 64-element tile
        std::vector<double> dt, ramp_duration;  // per instance
    } ens;
};

constexpr int kEnsembleTile = 256;  // instance padding: any block size of 64, 128, 256 stays instance-pure

// Builds the model; returns an empty string or an error message.
// reference_face_order: face slots in the reference's rhs accumulation order (the FP64 form);
// otherwise slots are the local faces 0, 1, 2 (the FP32 form).
std::string build_model(const rkdg_mesh_desc& md, const rkdg_params& prm, Model& M, bool reference_face_order = true);

// Merges models of equal polynomial order, RK scheme and physics options into one element
// range (an ensemble); returns an empty string or an error message.
std::string merge_models(const std::vector<const Model*>& parts, Model& out);

// L2 residual against the problem's true solution (element.hpp:547-592), FP64, host.
double l2_residual(const Model& M, const std::vector<double>& q, double t);

}  // namespace rkdg
