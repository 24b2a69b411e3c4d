// Per-element RKDG stage arithmetic shared by the HIP kernels.
//
// One work-item owns one element.  It rebuilds the stage right-hand side of that element
// from HBM-resident modal data (volume, source, the three face fluxes, M^-1), applies the
// SSP-RK combination and the wetting-drying post-step, and writes the next stage state.
// Interface fluxes are evaluated by both neighbours with identical operands and operation
// order, so each side reproduces the reference's single-evaluation value F_hat_in /
// F_hat_ex = -F_hat_in exactly (rkdg_swe_is_internal.hpp:35-45) without atomics or colouring.
//
// Arithmetic follows the reference term by term (sums over quadrature points and modes run
// left to right from the first product, Eigen 3.3 coefficient-product order; no FMA in the
// FP64 build), so the FP64 form can match the CPU oracle bit for bit where no libm
// transcendental is involved.  Faces are processed in the order the reference accumulates
// them into an element's rhs: interfaces by ascending neighbour ID (the (in, ex) map order
// of swe_pre_create_intface.hpp:14-31 seen from one element), then boundaries by BC type and
// face (swe_pre_create_bound.hpp:16-113).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RKDG_HD __host__ __device__
#else
#define RKDG_HD
#endif
// The per-element stage bodies are called from more than one kernel; they must inline into each
// (otherwise the kernel argument block is passed through scratch memory).
#define RKDG_FORCE_INLINE __attribute__((always_inline))


#include <cmath>

#include "device_tables.h"

namespace rkdg {

constexpr double kPI = 3.14159265359;  // general_definitions.hpp:227

template <int P>
struct Sizes {
    static constexpr int ND = (P + 1) * (P + 2) / 2;            // modes
    static constexpr int NG = P == 1 ? 3 : (P == 2 ? 6 : 12);   // Dunavant rule 2P points
    static constexpr int NE = P + 1;                            // Gauss-Legendre rule 2P+1 points
};

// Everything one stage launch reads and writes.  Arrays are element-fastest SoA.
template <typename Real>
struct StageArgs {
    int ne;                  // elements
    int nedge;               // edge records
    const Real* q_s;         // [3][ND][ne] state of the current stage
    Real* q_out;             // [3][ND][ne] next stage state
    Real* sol_out;           // [3][ND][ne] M^-1 rhs of this stage, when a later stage reads it
    int nterm;               // stored earlier-stage terms of the SSP-RK combination
    const Real* term_q[8];
    const Real* term_sol[8];
    Real term_a[8];
    Real term_c[8];          // dt * brk
    Real cur_a, cur_c;       // coefficients of the current stage's own q and sol
    const Real* bath;        // [ND][ne] modal bathymetry
    const Real* geom;        // [5][ne] Jinv(0,0) Jinv(0,1) Jinv(1,0) Jinv(1,1) |det J|
    const int* nb;           // [3][ne] per face slot: neighbour element, or -1-bc for a boundary
    const int* edge;         // [3][ne] per face slot: edge record
    const uint32_t* meta;    // [ne] per slot s: own face at bits 4s..4s+1, neighbour face at 4s+2..4s+3
    const Real* egeom;       // [3][nedge] nx, ny, surface J of the edge (in-side normal)
    const uint8_t* wet_in;   // [ne]
    uint8_t* wet_out;        // [ne]
    const double* bath_vrtx; // [3][ne] wetting-drying vertex depths (FP64 in every form)
    const Real* gmn2;        // [ne] g * mean(n)^2 (Manning)
    const Real* vxy;         // [6][ne] vertex x0 x1 x2 y0 y1 y2
    const int* eforce;       // [nedge] tide/flow forcing record, levee record, or -1
    const Real* levee_gp;    // [nlevee][3][NE] barrier height, C_subcritical, C_supercritical at in-side points
    const Real* sp_vol;      // [NG][ne] spherical correction at the volume points, or null (sp = 1; read
    const Real* sp_edge;     // [NE][nedge] ... and at the edge points (in-side order)  only by CF kernels)
    const Real* famp;        // [nforced][ncon][NE]
    const Real* fphase;      // [nforced][ncon][NE]
    int ncon;
    const Real* con;         // [ncon][3] frequency, forcing factor, equilibrium argument of each constituent
    double t;                // stage time
    double ramp;
    Real g;
    Real cf;                 // friction coefficient / Manning floor
    int friction;            // 0 none, 1 Chezy, 2 Manning
    int fsrc;                // function source
    int problem;
    double pparam[4];
    double h_o, h_thr;
    int nan_check;           // scan the new state for NaN here (no limiter after this stage)
    unsigned long long* nan_key;   // min over NaN finds of (stage_key << 32 | element)
    unsigned long long stage_key;  // step * nstages + stage of this launch (the stage alone with step_ctr)
    // Ensembles: several meshes merged into one element range, each padded to whole 256-element
    // tiles, with per-instance time, ramp and SSP-RK coefficients.  All null for a single mesh.
    const int* tile_inst;    // [ne/64] instance of each 64-element tile
    const int* inst_end;     // [ninst] one past the instance's last element
    const double* inst_tr;   // [ninst][2] stage time and ramp of each instance
    const Real* inst_coef;   // [ninst][9] cur_c, term_c[0..7] (dt * brk) of this stage per instance
    const unsigned long long* step_ctr;  // device step counter: stage_key = step * key_ns + stage_key
    int key_ns;
};

// Per-element view of the instance-dependent stage scalars.  ENS (compile time): an ensemble,
// whose time, ramp and dt * brk coefficients come from the per-instance tables; otherwise the
// launch's own.  Never chosen at run time: a run-time choice between a table and the kernel's
// argument block lets the compiler load through a selected pointer, which forces the whole
// argument block into scratch memory.
template <typename Real>
struct InstScal {
    double t, ramp;
    const Real* coef;  // ensembles: cur_c, term_c[0..7] of the element's instance
};

template <bool ENS, typename Real>
RKDG_HD inline InstScal<Real> inst_scal(const StageArgs<Real>& A, int inst) {
    if constexpr (ENS)
        return InstScal<Real>{A.inst_tr[2 * inst], A.inst_tr[2 * inst + 1], A.inst_coef + 9 * (size_t)inst};
    else
        return InstScal<Real>{A.t, A.ramp, nullptr};
}
template <bool ENS, typename Real>
RKDG_HD inline Real cur_coef(const StageArgs<Real>& A, const InstScal<Real>& IS) {
    if constexpr (ENS)
        return IS.coef[0];
    else
        return A.cur_c;
}
template <bool ENS, typename Real>
RKDG_HD inline Real term_coef(const StageArgs<Real>& A, const InstScal<Real>& IS, int i) {
    if constexpr (ENS)
        return IS.coef[1 + i];
    else
        return A.term_c[i];
}

// ---------------------------------------------------------------------------------------
// Element-fastest arrays: entry i of column c sits at base[c * n + i].  On the GPU a column
// access is a buffer instruction whose column offset lives in a scalar register, so a gather
// costs one instruction and no per-lane 64-bit address arithmetic.  Byte offsets stay below
// 2^32 (the model builder rejects larger meshes).

#ifdef __HIP_DEVICE_COMPILE__
RKDG_HD inline __amdgpu_buffer_rsrc_t rkdg_rsrc(const void* base) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, -1, 0x00020000);
}
#endif

template <typename T>
RKDG_HD inline T ldc(const T* base, int n, int c, int i) {
#ifdef __HIP_DEVICE_COMPILE__
    const unsigned soff = (unsigned)c * (unsigned)n * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rkdg_rsrc(base), i * 4, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rkdg_rsrc(base), i * 8, soff, 0));
#else
    return base[(size_t)c * n + i];
#endif
}

// Columns c = 0..ncol-1 of an element-fastest array for the 64 consecutive elements e0..e0+63 of
// a wave, copied HBM -> LDS by direct-to-LDS buffer loads (no registers held while they are in
// flight): column c lands as 64 values at dst[64 c ..].  Every lane of the wave must execute it.
// FP64: lane j fetches dword j of each 256-byte half of the column slice; FP32: lane j fetches
// element e0 + j.  Elements past n - 1 read element n - 1 (their values are never used).  The
// data may be read only after lds_columns_wait().
RKDG_HD inline void lds_columns_fetch(const double* base, int n, int ncol, int e0, double* dst) {
#ifdef __HIP_DEVICE_COMPILE__
    const int lane = (int)__lane_id();
    const int ea = e0 + (lane >> 1) < n ? e0 + (lane >> 1) : n - 1;
    const int eb = e0 + 32 + (lane >> 1) < n ? e0 + 32 + (lane >> 1) : n - 1;
    const int va = ea * 8 + (lane & 1) * 4, vb = eb * 8 + (lane & 1) * 4;
    const __amdgpu_buffer_rsrc_t r = rkdg_rsrc(base);
    for (int c = 0; c < ncol; ++c) {
        const int soff = (int)((unsigned)c * (unsigned)n * 8u);  // below 2^32 (model builder)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)(dst + 64 * c), 4, va,
                                                 soff, 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)(dst + 64 * c + 32), 4,
                                                 vb, soff, 0, 0);
    }
#else
    (void)base, (void)n, (void)ncol, (void)e0, (void)dst;
#endif
}

RKDG_HD inline void lds_columns_fetch(const float* base, int n, int ncol, int e0, float* dst) {
#ifdef __HIP_DEVICE_COMPILE__
    const int lane = (int)__lane_id();
    const int ea = e0 + lane < n ? e0 + lane : n - 1;
    const __amdgpu_buffer_rsrc_t r = rkdg_rsrc(base);
    for (int c = 0; c < ncol; ++c) {
        const int soff = (int)((unsigned)c * (unsigned)n * 4u);  // below 2^32 (model builder)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)(dst + 64 * c), 4, ea * 4,
                                                 soff, 0, 0);
    }
#else
    (void)base, (void)n, (void)ncol, (void)e0, (void)dst;
#endif
}

RKDG_HD inline int lane_index() {
#ifdef __HIP_DEVICE_COMPILE__
    return (int)__lane_id();
#else
    return 0;
#endif
}

// Waits for every outstanding vector-memory load of the wave, the direct-to-LDS ones included.
RKDG_HD inline void lds_columns_wait() {
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
#endif
}

template <typename T>
RKDG_HD inline void stc(T* base, int n, int c, int i, T v) {
#ifdef __HIP_DEVICE_COMPILE__
    const unsigned soff = (unsigned)c * (unsigned)n * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rkdg_rsrc(base), i * 4, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v),
                                              rkdg_rsrc(base), i * 8, soff, 0);
#else
    base[(size_t)c * n + i] = v;
#endif
}

// ---------------------------------------------------------------------------------------
// Closed-form problem functions (examples/swe_manufactured_solution/*.hpp,
// examples/swe_parabolic_bowl/parabolic_bowl_initial_condition_functions.hpp:7-28).

template <typename Real>
RKDG_HD inline void problem_ic(int problem, const double* pp, Real g, double t, Real x, Real y, Real* q) {
    using std::cos;
    using std::exp;
    using std::sin;
    if (problem == 1) {
        q[0] = exp(sin(3 * x) * sin(3 * y) - sin(Real(3 * t)));
        q[1] = cos(x - Real(4 * t));
        q[2] = sin(y + Real(4 * t));
    } else if (problem == 2) {
        const Real alpha = Real(1.e-7), X = 1, Y = Real(-0.41884);
        Real r2 = x * x + y * y;
        Real bathymetry = -alpha * r2;
        Real H = 1 / (X + Y) + alpha * (Y * Y - X * X) / ((X + Y) * (X + Y)) * r2;
        q[0] = H - bathymetry, q[1] = 0, q[2] = 0;
    } else if (problem == 3) {
        q[0] = x < Real(pp[0]) ? Real(pp[1]) : Real(pp[2]);
        q[1] = 0, q[2] = 0;
    } else {
        q[0] = 0, q[1] = 0, q[2] = 0;
    }
}

template <typename Real>
RKDG_HD inline void problem_source(double td, Real x, Real y, Real* s) {
    using std::cos;
    using std::exp;
    using std::pow;
    using std::sin;
    const Real t = Real(td);
    s[0] = -3 * exp(-sin(3 * t) + sin(3 * x) * sin(3 * y)) * cos(3 * t) + cos(4 * t + y) + sin(4 * t - x);
    const Real E = exp(-sin(3 * t) + sin(3 * x) * sin(3 * y));
    s[1] = (cos(4 * t - x) * cos(4 * t + y)) / (2 + E) - 4 * sin(4 * t - x) +
           (2 * cos(4 * t - x) * sin(4 * t - x)) / (2 + E) + Real(3.) * E * (2 + E) * cos(3 * x) * sin(3 * y) -
           (3 * E * (cos(4 * t - x) * cos(4 * t - x)) * cos(3 * x) * sin(3 * y)) / ((2 + E) * (2 + E)) -
           (3 * E * cos(4 * t - x) * cos(3 * y) * sin(3 * x) * sin(4 * t + y)) / ((2 + E) * (2 + E));
    s[2] = 4 * cos(4 * t + y) + Real(3.) * E * (2 + E) * cos(3 * y) * sin(3 * x) +
           (2 * cos(4 * t + y) * sin(4 * t + y)) / (2 + E) + (sin(4 * t - x) * sin(4 * t + y)) / (2 + E) -
           (3 * E * cos(4 * t - x) * cos(3 * x) * sin(3 * y) * sin(4 * t + y)) / ((2 + E) * (2 + E)) -
           (3 * E * cos(3 * y) * sin(3 * x) * (sin(4 * t + y) * sin(4 * t + y))) / ((2 + E) * (2 + E));
}

// ---------------------------------------------------------------------------------------
// Local Lax-Friedrichs flux (numerical_fluxes/rkdg_swe_LLF_flux.hpp:7-60).  bath and sp (the
// spherical correction, 1 on a Cartesian mesh) are those of the side whose aux the caller passes;
// sp scales the in-side flux's x-terms and the wave speed's x-part, not the ex-side flux (:45-57).

template <typename Real>
RKDG_HD inline Real rmax(Real a, Real b) {  // std::max semantics: (a < b) ? b : a
    return (a < b) ? b : a;
}

// Depth under the wave speed's square root.  FP32: a depth the FP64 wetting-drying pass left at
// or just above zero at a vertex can come back from the FP32 modes a rounding error below zero at
// a trace point; it is taken as zero there (the square root would give NaN).  FP64: unchanged.
RKDG_HD inline float depth_for_speed(float h) { return h > 0.0f ? h : 0.0f; }
RKDG_HD inline double depth_for_speed(double h) { return h; }

// Depth under a velocity's division (u = qx / h).  FP32: a dry element is flattened by the FP64
// wetting-drying pass to the depth floor thr (about 1e-12) with zero momentum, and its depth comes
// back from the FP32 modes within rounding of zero (ze of metres carries 1e-7), or below; it is
// taken as the smallest normal float there, so the zero momentum gives a zero velocity (0 / 0
// would be NaN).  FP64 keeps such depths positive (rounding 1e-16 against thr): unchanged.
RKDG_HD inline float depth_for_div(float h) { return h > 1.17549435e-38f ? h : 1.17549435e-38f; }
RKDG_HD inline double depth_for_div(double h) { return h; }
template <typename Real>
RKDG_HD inline Real rmin(Real a, Real b) {  // std::min semantics: (b < a) ? b : a
    return (b < a) ? b : a;
}

template <typename Real>
RKDG_HD inline void llf(Real grav, const Real* qi, const Real* qx, Real bath, Real nx, Real ny, Real* F,
                        Real sp = Real(1)) {
    using std::fabs;
    using std::sqrt;
    const Real h_in = qi[0] + bath;
    const Real u_in = qi[1] / depth_for_div(h_in);
    const Real v_in = qi[2] / depth_for_div(h_in);
    const Real h_ex = qx[0] + bath;
    const Real u_ex = qx[1] / depth_for_div(h_ex);
    const Real v_ex = qx[2] / depth_for_div(h_ex);
    const Real un_in = u_in * nx + v_in * ny;
    const Real un_ex = u_ex * nx + v_ex * ny;
    const Real spc = (nx * sp) * (nx * sp) + ny * ny;
    const Real lam = rmax(fabs(un_in) + sqrt(grav * depth_for_speed(h_in) * spc),
                          fabs(un_ex) + sqrt(grav * depth_for_speed(h_ex) * spc));
    const Real uuh_in = u_in * qi[1], vvh_in = v_in * qi[2], uvh_in = u_in * qi[2];
    const Real pe_in = grav * ((qi[0] * qi[0]) / 2 + qi[0] * bath);
    const Real uuh_ex = u_ex * qx[1], vvh_ex = v_ex * qx[2], uvh_ex = u_ex * qx[2];
    const Real pe_ex = grav * ((qx[0] * qx[0]) / 2 + qx[0] * bath);
    const Real fi0 = sp * qi[1] * nx + qi[2] * ny;
    const Real fi1 = sp * (uuh_in + pe_in) * nx + uvh_in * ny;
    const Real fi2 = sp * uvh_in * nx + (vvh_in + pe_in) * ny;
    const Real fe0 = qx[1] * nx + qx[2] * ny;
    const Real fe1 = (uuh_ex + pe_ex) * nx + uvh_ex * ny;
    const Real fe2 = uvh_ex * nx + (vvh_ex + pe_ex) * ny;
    F[0] = Real(0.5) * (fi0 + fe0 + lam * (qi[0] - qx[0]));
    F[1] = Real(0.5) * (fi1 + fe1 + lam * (qi[1] - qx[1]));
    F[2] = Real(0.5) * (fi2 + fe2 + lam * (qi[2] - qx[2]));
}

// Reflective (land) exterior state + LLF (boundary_conditions/rkdg_swe_bc_land.hpp:68-86).
template <typename Real>
RKDG_HD inline void land_llf(Real grav, Real nx, Real ny, const Real* qi, Real bath, Real* F, Real sp = Real(1)) {
    const Real tx = -ny, ty = nx;
    const Real qn_ex = -(qi[1] * nx + qi[2] * ny);
    const Real qt_ex = qi[1] * tx + qi[2] * ty;
    Real qe[3] = {qi[0], qn_ex * nx + qt_ex * tx, qn_ex * ny + qt_ex * ty};
    llf(grav, qi, qe, bath, nx, ny, F, sp);
}

// Interface flux seen by one side (ISP::Internal::ComputeFlux, rkdg_swe_is_internal.hpp:24-94).
// qi/qx: in- and ex-side traces at one in-side point and its mirrored ex-side point;
// bi/bx: their bathymetry.  Returns the flux of the requested side.
template <typename Real, bool WD>
RKDG_HD inline void interface_flux(Real grav, const Real* qi, const Real* qx, Real bi, Real bx, Real nx, Real ny,
                                   bool wet_in, bool wet_ex, bool want_in, Real* F, Real sp = Real(1)) {
    Real Fi[3], Fe[3];
    llf(grav, qi, qx, bi, nx, ny, Fi, sp);
    Fe[0] = -Fi[0], Fe[1] = -Fi[1], Fe[2] = -Fi[2];
    if (WD) {
        if (Fi[0] > Real(1e-12)) {
            if (!wet_in) {
                Fi[0] = 0, Fi[1] = 0, Fi[2] = 0;
                land_llf(grav, -nx, -ny, qx, bx, Fe, sp);
            } else if (!wet_ex) {
                const Real f0 = Fi[0];
                llf(Real(0), qx, qi, bx, -nx, -ny, Fe, sp);
                Fe[0] = -f0;
            }
        } else if (Fi[0] < Real(-1e-12)) {
            if (!wet_ex) {
                Fe[0] = 0, Fe[1] = 0, Fe[2] = 0;
                land_llf(grav, nx, ny, qi, bi, Fi, sp);
            } else if (!wet_in) {
                const Real fe0 = Fe[0];
                llf(Real(0), qi, qx, bi, nx, ny, Fi, sp);
                Fi[0] = -fe0;
            }
        }
    }
    const Real* src = want_in ? Fi : Fe;
    F[0] = src[0], F[1] = src[1], F[2] = src[2];
}

// Levee (internal barrier) flux of one side, ISP::Levee::ComputeFlux
// (interface_specializations/rkdg_swe_is_levee.hpp:62-174) seen from that side.  qo: this
// side's trace at its point, qnb: the other side's at the mirrored point, bo: this side's
// bathymetry, (nx, ny): the in-side normal (the ex side's is its negation), H/Cs/Cp: barrier
// height and weir coefficients at the in-side point.  Both sides below the crest, or level
// within 1 cm: reflective (Land) exterior state.  Otherwise the higher side overtops: the
// exterior momentum is the weir discharge (subcritical when the lower side's head exceeds
// 2/3 of the higher's) along the overtopping side's normal; a dry receiving side sees no
// gravity in its LLF flux.
template <typename Real>
RKDG_HD inline void levee_flux(Real grav, bool is_in, const Real* qo, const Real* qnb, Real bo, Real nx, Real ny,
                               Real H, Real Cs, Real Cp, bool wet_own, Real* F, Real sp = Real(1)) {
    using std::abs;
    using std::sqrt;
    const Real tol = Real(0.01);  // H_tolerance
    const Real ze_in = is_in ? qo[0] : qnb[0], ze_ex = is_in ? qnb[0] : qo[0];
    const Real h_in = ze_in - H, h_ex = ze_ex - H;
    const Real onx = is_in ? nx : -nx, ony = is_in ? ny : -ny;  // this side's normal
    Real qe[3], gravity = grav;
    if ((h_in <= tol && h_ex <= tol) || abs(h_in - h_ex) <= tol) {  // reflective on both sides
        const Real tx = -ony, ty = onx;
        const Real qn_ex = -(qo[1] * onx + qo[2] * ony);
        const Real qt_ex = qo[1] * tx + qo[2] * ty;
        qe[0] = qo[0], qe[1] = qn_ex * onx + qt_ex * tx, qe[2] = qn_ex * ony + qt_ex * ty;
    } else {
        const bool from_in = h_in > h_ex;  // overtopping from in to ex, else from ex to in
        const Real hu = from_in ? h_in : h_ex, hd = from_in ? h_ex : h_in;  // upstream / downstream heads
        const Real n_x = from_in ? nx : -nx, n_y = from_in ? ny : -ny;
        const Real t_x = -n_y, t_y = n_x;
        Real qn;
        if (hd > Real(2.0) * hu / Real(3.0))  // subcritical
            qn = Cs * hd * sqrt(Real(2.0) * grav * (hu - hd));
        else  // supercritical
            qn = Cp * (Real(2.0) * hu / Real(3.0)) * sqrt(grav * (Real(2.0) * hu / Real(3.0)));
        const Real qt = Real(0.0);
        qe[0] = qo[0], qe[1] = qn * n_x + qt * t_x, qe[2] = qn * n_y + qt * t_y;
        const bool receiving = from_in ? !is_in : is_in;
        if (receiving && !wet_own) gravity = Real(0.0);
    }
    llf(gravity, qo, qe, bo, onx, ony, F, sp);
}

// ---------------------------------------------------------------------------------------
// Wetting-drying of one element on its linear (vertex) state, in FP64
// (problem_postprocessor/swe_post_wet_dry.hpp:5-145).  qv[var][vertex] in/out; returns the
// new wet flag; *rewrite says whether the modal state must be rebuilt from qv.  tol widens the
// comparisons with h_o (dry vertex, flattening, re-wetting) by the rounding of the stored state:
// 0 for FP64 states, which reproduce the reference exactly.
RKDG_HD inline bool wet_dry_vertices(double qv[3][3], const double* bv, double bmin, bool wet, double h_o,
                                     double thr, double tol, bool* rewrite) {
    *rewrite = false;
    double h[3];
    unsigned n_dry = 0;
    for (int v = 0; v < 3; ++v) {
        h[v] = qv[0][v] + bv[v];
        if (h[v] <= h_o + thr + tol) ++n_dry;
    }
    double h_avg = (((0.0 + h[0]) + h[1]) + h[2]) / 3;
    if (h_avg <= thr) h_avg = thr;
    bool check = false;
    if (n_dry == 0) {
        if (wet) return true;
        check = true;
    } else if (n_dry == 3) {
        for (int v = 0; v < 3; ++v) qv[0][v] = h_avg - bv[v], qv[1][v] = 0.0, qv[2][v] = 0.0;
        *rewrite = true;
        return false;
    } else {
        if (h_avg <= h_o + thr + tol) {
            for (int v = 0; v < 3; ++v) qv[0][v] = h_avg - bv[v], qv[1][v] = 0, qv[2][v] = 0;
        } else {
            int hmin = 0, hmax = 0;
            for (int v = 1; v < 3; ++v) {
                if (h[v] < h[hmin]) hmin = v;
                if (h[hmax] < h[v]) hmax = v;
            }
            const int hmid = 3 - hmax - hmin;
            double ht[3];
            ht[hmin] = h_o;
            ht[hmid] = rmax(h_o, h[hmid] - (ht[hmin] - h[hmin]) / 2);
            ht[hmax] = h[hmax] - (ht[hmin] - h[hmin]) - (ht[hmid] - h[hmid]);
            for (int v = 0; v < 3; ++v) h[v] = ht[v];
            double dqx = 0, dqy = 0;
            unsigned nd = 0;
            for (int v = 0; v < 3; ++v)
                if (h[v] <= h_o + thr + tol) {
                    nd++;
                    dqx += qv[1][v], dqy += qv[2][v];
                    qv[1][v] = 0.0, qv[2][v] = 0.0;
                }
            for (int v = 0; v < 3; ++v) {
                qv[0][v] = h[v] - bv[v];
                if (h[v] > h_o + thr + tol) {
                    qv[1][v] += dqx / (3 - nd);
                    qv[2][v] += dqy / (3 - nd);
                }
            }
        }
        *rewrite = true;
        check = true;
    }
    bool new_wet = wet;
    if (check) {
        int hmax = 0;
        for (int v = 1; v < 3; ++v)
            if (h[hmax] < h[v]) hmax = v;
        if (qv[0][hmax] > h_o - bmin + tol) {
            new_wet = true;
        } else {
            new_wet = false;
            for (int v = 0; v < 3; ++v) qv[1][v] = 0.0, qv[2][v] = 0.0;
            *rewrite = true;
        }
    }
    return new_wet;
}

// Modal -> linear vertex values, q * T_basis_linear (basis_dubiner_2D.cpp:213-234).
template <typename Real>
RKDG_HD inline void to_vertices(const Real* q, Real* v) {
    v[0] = (q[0] * Real(1.0) + q[1] * Real(-1.0)) + q[2] * Real(-1.0);
    v[1] = (q[0] * Real(1.0) + q[1] * Real(-1.0)) + q[2] * Real(1.0);
    v[2] = (q[0] * Real(1.0) + q[1] * Real(2.0)) + q[2] * Real(0.0);
}
// Vertex values -> modes 0..2, q_lin * T_linear_basis (basis_dubiner_2D.cpp:236-257).
template <typename Real>
RKDG_HD inline void from_vertices(const Real* v, Real* q) {
    q[0] = (v[0] * Real(1.0 / 3.0) + v[1] * Real(1.0 / 3.0)) + v[2] * Real(1.0 / 3.0);
    q[1] = (v[0] * Real(-1.0 / 6.0) + v[1] * Real(-1.0 / 6.0)) + v[2] * Real(1.0 / 3.0);
    q[2] = (v[0] * Real(-1.0 / 2.0) + v[1] * Real(1.0 / 2.0)) + v[2] * Real(0.0);
}

// ---------------------------------------------------------------------------------------
// Face tables indexed by a per-lane local face: phie[(f*ND + d)*NE + g], psie[(f*3 + node)*NE + g].
// The kernels stage them in LDS; the remaining reference-element tables are compile-time
// constants (ConstTables, generated from tables.cpp at build time).
template <typename Real>
struct FaceTables {
    const Real* phie;
    const Real* psie;
};

// Tide/flow exterior forcing value at one edge point (rkdg_swe_bc_tide.hpp:76-83).
template <typename Real, int NE>
RKDG_HD inline Real harmonic_forcing(const StageArgs<Real>& A, int ed, int g, double t, double ramp) {
    const int fr = A.eforce[ed];
    Real acc = Real(0.0);
    for (int c = 0; c < A.ncon; ++c) {
        using std::cos;
        const Real amp = A.famp[((size_t)fr * A.ncon + c) * NE + g];
        const Real ph = A.fphase[((size_t)fr * A.ncon + c) * NE + g];
        const Real freq = A.con[3 * c], ffact = A.con[3 * c + 1], eqarg = A.con[3 * c + 2];
        acc += Real(ramp) * ffact * amp * cos(freq * Real(t) + (eqarg - ph) * Real(kPI) / Real(180));
    }
    return acc;
}

// Wetting-drying of the new state (swe_post_wet_dry.hpp:5-145, FP64), store, NaN scan.
template <typename Real, int P, bool WD>
RKDG_HD RKDG_FORCE_INLINE inline void finish_element(const StageArgs<Real>& A, int e, bool wet, Real (&nq)[3][Sizes<P>::ND]) {
    constexpr int ND = Sizes<P>::ND;
    const int ne = A.ne;
    if (WD) {
        double qv[3][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double m[3] = {(double)nq[k][0], (double)nq[k][1], (double)nq[k][2]};
            to_vertices(m, qv[k]);
        }
        const double bv[3] = {ldc(A.bath_vrtx, ne, 0, e), ldc(A.bath_vrtx, ne, 1, e), ldc(A.bath_vrtx, ne, 2, e)};
        const double bmin = rmin(rmin(bv[0], bv[1]), bv[2]);
        // FP32 state: the vertex elevations come back from FP32 modes with a rounding of about
        // 2^-23 of their size (1e-7 for elevations of metres), far above the reference's depth
        // tolerance h_thr (1e-12 of a metre).  An element flattened at h_o would otherwise read as
        // partly wet on one side of h_o, be taken as wet, and keep the momentum its wet
        // neighbours' face fluxes gave it while dry (tens of metres per second in a centimetre
        // of water): the comparisons with h_o allow 8 units of that rounding.
        double tol = 0.0;
        if constexpr (sizeof(Real) == 4)
            tol = 0x1p-20 * rmax(rmax(fabs(qv[0][0]), fabs(qv[0][1])), fabs(qv[0][2]));
        bool rewrite = false;
        const bool wet_new = wet_dry_vertices(qv, bv, bmin, wet, A.h_o, A.h_thr, tol, &rewrite);
        if (rewrite)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double m[3];
                from_vertices(qv[k], m);
#pragma unroll
                for (int d = 0; d < 3; ++d) nq[k][d] = (Real)m[d];
#pragma unroll
                for (int d = 3; d < ND; ++d) nq[k][d] = Real(0.0);
            }
        A.wet_out[e] = wet_new ? 1 : 0;
    }
    bool has_nan = false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            stc(A.q_out, ne, k * ND + d, e, nq[k][d]);
            has_nan |= (nq[k][d] != nq[k][d]);
        }
    if (A.nan_check && has_nan) {
        const unsigned long long sk = A.step_ctr ? *A.step_ctr * (unsigned long long)A.key_ns + A.stage_key : A.stage_key;
        const unsigned long long key = (sk << 32) | (unsigned long long)(unsigned)e;
#ifdef __HIP_DEVICE_COMPILE__
        atomicMin(A.nan_key, key);
#else
        if (key < *A.nan_key) *A.nan_key = key;
#endif
    }
}

// ---------------------------------------------------------------------------------------
// The fused stage of one element, reference arithmetic order (FP64 plain form).

template <typename Real, int P, bool WD, bool CF, bool ENS>
RKDG_HD RKDG_FORCE_INLINE inline void stage_element(const StageArgs<Real>& A, const FaceTables<Real>& FT, int e, const InstScal<Real>& IS,
                                                    const Real* t0_lds) {
    using C = ConstTables<Real, P>;
    constexpr int ND = Sizes<P>::ND, NG = Sizes<P>::NG, NE = Sizes<P>::NE;
    const int ne = A.ne;
    const bool wet = WD ? (A.wet_in[e] != 0) : true;
    const Real grav = A.g;

    Real q[3][ND];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) q[k][d] = ldc(A.q_s, ne, k * ND + d, e);
    Real bath[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) bath[d] = ldc(A.bath, ne, d, e);
    const Real J00 = ldc(A.geom, ne, 0, e), J01 = ldc(A.geom, ne, 1, e), J10 = ldc(A.geom, ne, 2, e),
               J11 = ldc(A.geom, ne, 3, e), adet = ldc(A.geom, ne, 4, e);

    Real rhs[3][ND];
    if (wet) {
        // ---- volume_kernel (rkdg_swe_proc_volume.hpp:8-35): rhs = Fx*Dx + Fy*Dy
        Real ax[3][ND], ay[3][ND];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            Real qg[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Real s = q[k][0] * C::phi[0][g];
#pragma unroll
                for (int d = 1; d < ND; ++d) s = s + q[k][d] * C::phi[d][g];
                qg[k] = s;
            }
            Real bg = bath[0] * C::phi[0][g];
#pragma unroll
            for (int d = 1; d < ND; ++d) bg = bg + bath[d] * C::phi[d][g];
            const Real h = qg[0] + bg;
            const Real u = qg[1] / h, v = qg[2] / h;
            const Real uuh = u * qg[1], vvh = v * qg[2], uvh = u * qg[2];
            const Real pe = grav * (Real(0.5) * (qg[0] * qg[0]) + qg[0] * bg);
            Real fx[3] = {qg[1], uuh + pe, uvh};
            const Real fy[3] = {qg[2], uvh, vvh + pe};
            if (CF && A.sp_vol) {  // Fx *= sp (rkdg_swe_proc_volume.hpp:25-30)
                const Real sp = ldc(A.sp_vol, ne, g, e);
                fx[0] = sp * fx[0], fx[1] = sp * fx[1], fx[2] = sp * fx[2];
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real dx = Real(0.0);
                dx += C::idphi0[g][d] * J00;
                dx += C::idphi1[g][d] * J10;
                dx *= adet;
                Real dy = Real(0.0);
                dy += C::idphi0[g][d] * J01;
                dy += C::idphi1[g][d] * J11;
                dy *= adet;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (g == 0) {
                        ax[k][d] = fx[k] * dx;
                        ay[k][d] = fy[k] * dy;
                    } else {
                        ax[k][d] = ax[k][d] + fx[k] * dx;
                        ay[k][d] = ay[k][d] + fy[k] * dy;
                    }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int d = 0; d < ND; ++d) rhs[k][d] = ax[k][d] + ay[k][d];

        // ---- source_kernel (rkdg_swe_proc_source.hpp:8-18, swe_source.hpp:8-97)
        Real ts[3][ND];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            Real qg[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Real s = q[k][0] * C::phi[0][g];
#pragma unroll
                for (int d = 1; d < ND; ++d) s = s + q[k][d] * C::phi[d][g];
                qg[k] = s;
            }
            Real dbx = Real(0), dby = Real(0);
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                Real ex = Real(0.0);
                ex += C::dphi0[d][g] * J00;
                ex += C::dphi1[d][g] * J10;
                Real ey = Real(0.0);
                ey += C::dphi0[d][g] * J01;
                ey += C::dphi1[d][g] * J11;
                if (d == 0) {
                    dbx = bath[0] * ex;
                    dby = bath[0] * ey;
                } else {
                    dbx = dbx + bath[d] * ex;
                    dby = dby + bath[d] * ey;
                }
            }
            Real src[3] = {Real(0.0), grav * (dbx * qg[0]), grav * (dby * qg[0])};
            if (CF && A.fsrc) {
                const Real x = (A.vxy[e] * C::psi[0][g] + A.vxy[(size_t)ne + e] * C::psi[1][g]) +
                               A.vxy[2 * (size_t)ne + e] * C::psi[2][g];
                const Real y = (A.vxy[3 * (size_t)ne + e] * C::psi[0][g] + A.vxy[4 * (size_t)ne + e] * C::psi[1][g]) +
                               A.vxy[5 * (size_t)ne + e] * C::psi[2][g];
                Real f[3];
                problem_source<Real>(IS.t, x, y, f);
#pragma unroll
                for (int k = 0; k < 3; ++k) src[k] = src[k] + f[k];
            }
            if (A.friction) {
                Real bg = bath[0] * C::phi[0][g];
#pragma unroll
                for (int d = 1; d < ND; ++d) bg = bg + bath[d] * C::phi[d][g];
                const Real h = qg[0] + bg;
                const Real u = qg[1] / h, v = qg[2] / h;
                Real Cf = A.cf;
                if (A.friction == 2) {
                    using std::pow;
                    Cf = ldc(A.gmn2, ne, 0, e) / pow(h, Real(1.0 / 3.0));
                    if (Cf < A.cf) Cf = A.cf;
                }
                using std::hypot;
                const Real tau = Cf * hypot(u, v) / h;
                src[1] -= tau * qg[1];
                src[2] -= tau * qg[2];
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const Real w = C::ipf[g][d] * adet;
#pragma unroll
                for (int k = 0; k < 3; ++k) ts[k][d] = (g == 0) ? src[k] * w : ts[k][d] + src[k] * w;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int d = 0; d < ND; ++d) rhs[k][d] = rhs[k][d] + ts[k][d];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int d = 0; d < ND; ++d) rhs[k][d] = Real(0.0);
    }

    // ---- faces: interface_kernel / boundary_kernel contributions, in reference order
    const uint32_t meta = ldc(A.meta, ne, 0, e);
#pragma unroll 1
    for (int s = 0; s < 3; ++s) {
        const int nbv = ldc(A.nb, ne, s, e);
        const int ed = ldc(A.edge, ne, s, e);
        const int f = (meta >> (4 * s)) & 3;
        const int fn = (meta >> (4 * s + 2)) & 3;
        const Real* pf = FT.phie + (size_t)f * ND * NE;
        const bool interior = nbv >= 0;
        const bool wet_nb = (WD && interior) ? (A.wet_in[nbv] != 0) : true;
        if (interior ? !(wet || wet_nb) : !wet) continue;
        const Real nx = ldc(A.egeom, A.nedge, 0, ed), ny = ldc(A.egeom, A.nedge, 1, ed), sJ = ldc(A.egeom, A.nedge, 2, ed);
        Real qo[3][NE], bo[NE];
#pragma unroll
        for (int g = 0; g < NE; ++g) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                Real acc = q[k][0] * pf[g];
#pragma unroll
                for (int d = 1; d < ND; ++d) acc = acc + q[k][d] * pf[d * NE + g];
                qo[k][g] = acc;
            }
            Real acc = bath[0] * pf[g];
#pragma unroll
            for (int d = 1; d < ND; ++d) acc = acc + bath[d] * pf[d * NE + g];
            bo[g] = acc;
        }
        Real F[3][NE];
        const bool lev = (meta >> (12 + s)) & 1;
        if (interior) {
            const bool is_in = e < nbv;
            const Real* pn = FT.phie + (size_t)fn * ND * NE;
            Real qn[3][NE], bn[NE];
            {
                Real qnb[3][ND], bnb[ND];
#pragma unroll
                for (int k = 0; k < 3; ++k)
#pragma unroll
                    for (int d = 0; d < ND; ++d) qnb[k][d] = ldc(A.q_s, ne, k * ND + d, nbv);
#pragma unroll
                for (int d = 0; d < ND; ++d) bnb[d] = ldc(A.bath, ne, d, nbv);
#pragma unroll
                for (int g = 0; g < NE; ++g) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        Real acc = qnb[k][0] * pn[g];
#pragma unroll
                        for (int d = 1; d < ND; ++d) acc = acc + qnb[k][d] * pn[d * NE + g];
                        qn[k][g] = acc;
                    }
                    Real acc = bnb[0] * pn[g];
#pragma unroll
                    for (int d = 1; d < ND; ++d) acc = acc + bnb[d] * pn[d * NE + g];
                    bn[g] = acc;
                }
            }
            if (lev) {  // levee interface (ISP::Levee)
                const Real* lp = A.levee_gp + (size_t)A.eforce[ed] * 3 * NE;
#pragma unroll
                for (int g = 0; g < NE; ++g) {
                    const int gm = NE - 1 - g, gi = is_in ? g : gm;  // gi: the in-side point
                    const Real qo_g[3] = {qo[0][g], qo[1][g], qo[2][g]};
                    const Real qn_g[3] = {qn[0][gm], qn[1][gm], qn[2][gm]};
                    Real Fg[3];
                    const Real sp = (CF && A.sp_edge) ? ldc(A.sp_edge, A.nedge, gi, ed) : Real(1);
                    levee_flux<Real>(grav, is_in, qo_g, qn_g, bo[g], nx, ny, lp[gi], lp[NE + gi], lp[2 * NE + gi], wet, Fg,
                                     sp);
                    F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
                }
            } else {
#pragma unroll
                for (int g = 0; g < NE; ++g) {
                    const int gm = NE - 1 - g;  // mirrored point on the other side
                    const Real sp = (CF && A.sp_edge) ? ldc(A.sp_edge, A.nedge, is_in ? g : gm, ed) : Real(1);
                    Real Fg[3];
                    if (is_in) {
                        const Real qi[3] = {qo[0][g], qo[1][g], qo[2][g]};
                        const Real qx[3] = {qn[0][gm], qn[1][gm], qn[2][gm]};
                        interface_flux<Real, WD>(grav, qi, qx, bo[g], bn[gm], nx, ny, wet, wet_nb, true, Fg, sp);
                    } else {
                        const Real qi[3] = {qn[0][gm], qn[1][gm], qn[2][gm]};
                        const Real qx[3] = {qo[0][g], qo[1][g], qo[2][g]};
                        interface_flux<Real, WD>(grav, qi, qx, bn[gm], bo[g], nx, ny, wet_nb, wet, false, Fg, sp);
                    }
                    F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
                }
            }
        } else {
            const int bc = -1 - nbv;
#pragma unroll
            for (int g = 0; g < NE; ++g) {
                const Real qi[3] = {qo[0][g], qo[1][g], qo[2][g]};
                Real qe[3];
                if (bc == 0) {  // Land (rkdg_swe_bc_land.hpp:40-66)
                    const Real tx = -ny, ty = nx;
                    const Real qn_ex = -(qi[1] * nx + qi[2] * ny);
                    const Real qt_ex = qi[1] * tx + qi[2] * ty;
                    qe[0] = qi[0], qe[1] = qn_ex * nx + qt_ex * tx, qe[2] = qn_ex * ny + qt_ex * ty;
                } else if (bc == 1 || bc == 2) {  // Tide / Flow (rkdg_swe_bc_tide.hpp:70-96, rkdg_swe_bc_flow.hpp:72-101)
                    const Real acc = harmonic_forcing<Real, NE>(A, ed, g, IS.t, IS.ramp);
                    if (bc == 1)
                        qe[0] = acc, qe[1] = qi[1], qe[2] = qi[2];
                    else
                        qe[0] = qi[0], qe[1] = acc * nx, qe[2] = acc * ny;
                } else if (CF) {  // Function (rkdg_swe_bc_function.hpp:139-172): ic_q(t, x_gp)
                    const Real* ps = FT.psie + (size_t)f * 3 * NE;
                    const Real x = (A.vxy[e] * ps[g] + A.vxy[(size_t)ne + e] * ps[NE + g]) +
                                   A.vxy[2 * (size_t)ne + e] * ps[2 * NE + g];
                    const Real y = (A.vxy[3 * (size_t)ne + e] * ps[g] + A.vxy[4 * (size_t)ne + e] * ps[NE + g]) +
                                   A.vxy[5 * (size_t)ne + e] * ps[2 * NE + g];
                    problem_ic<Real>(A.problem, A.pparam, grav, IS.t, x, y, qe);
                } else {
                    qe[0] = qi[0], qe[1] = qi[1], qe[2] = qi[2];  // unreachable: function BCs launch the CF build
                }
                Real Fg[3];
                llf(grav, qi, qe, bo[g], nx, ny, Fg, (CF && A.sp_edge) ? ldc(A.sp_edge, A.nedge, g, ed) : Real(1));
                F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
            }
        }
        // rhs -= F * (phi_e * (w sJ))   (IntegrationPhiIN/EX, interface.hpp:170-182)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            Real acc[3];
#pragma unroll
            for (int g = 0; g < NE; ++g) {
                const Real w = pf[d * NE + g] * (C::we[0][g] * sJ);
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[k] = (g == 0) ? F[k][g] * w : acc[k] + F[k][g] * w;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) rhs[k][d] = rhs[k][d] - acc[k];
        }
    }

    // ---- M^-1 and the SSP-RK combination (element.hpp:500-505, explicit_ssp_rk_stepper.hpp:67-80);
    // the first earlier state may wait in LDS (t0_lds, fetched when the stage started)
    if (t0_lds) lds_columns_wait();
    // the first term's scalars are read once, outside the mode loop (a kernel-argument load
    // inside it waits out its latency for every mode); further terms (schemes with more than two
    // stages behind) take the general loop
    const int nterm = A.nterm;
    const Real ta0 = A.term_a[0], tc0 = term_coef<ENS>(A, IS, 0), ca = A.cur_a, cc = cur_coef<ENS>(A, IS);
    const Real* const tq0 = A.term_q[0];
    const Real* const ts0 = A.term_sol[0];
    Real* const so = A.sol_out;
    Real nq[3][ND];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const Real sol = rhs[k][d] * (C::minv[0][d] / adet);
            if (so) stc(so, ne, k * ND + d, e, sol);
            Real acc = Real(0.0);
            if (nterm > 0) {
                Real term = ta0 * (t0_lds ? t0_lds[64 * (k * ND + d) + lane_index()] : ldc(tq0, ne, k * ND + d, e));
                if (ts0) term = term + tc0 * ldc(ts0, ne, k * ND + d, e);
                acc = acc + term;
                for (int i = 1; i < nterm; ++i) {
                    term = A.term_a[i] * ldc(A.term_q[i], ne, k * ND + d, e);
                    if (A.term_sol[i]) term = term + term_coef<ENS>(A, IS, i) * ldc(A.term_sol[i], ne, k * ND + d, e);
                    acc = acc + term;
                }
            }
            acc = acc + (ca * q[k][d] + cc * sol);
            nq[k][d] = acc;
        }
    finish_element<Real, P, WD>(A, e, wet, nq);
}

// ---------------------------------------------------------------------------------------
// FP32 bulk form: the same stage with fewer operations.  Fluxes are rotated to reference
// coordinates once per point (G1 = J00 Fx + J01 Fy, G2 = J10 Fx + J11 Fy) instead of building the
// per-element D matrices; volume, source and lifts accumulate unscaled and M^-1/|det J| is applied
// once; divisions become reciprocals; faces run in local-face order so the own-face tables are
// compile-time constants (the neighbour's face table comes from LDS).  Both neighbours still
// evaluate each interface flux, each from its own side; the mass flux stays exactly
// antisymmetric, so mass is conserved to roundoff.

// Fast-form elementary operations.  FP32: hardware reciprocal / square root / exp2-log2
// (about 1 ulp).  FP64: hardware reciprocal refined by two Newton steps, square root from the
// hardware reciprocal square root refined by one coupled Newton step and a final residual
// correction, and x^(-1/3) from an FP32 estimate refined by two Newton steps.
RKDG_HD inline float fmaR(float a, float b, float c) { return fmaf(a, b, c); }
RKDG_HD inline double fmaR(double a, double b, double c) { return fma(a, b, c); }
RKDG_HD inline float absR(float a) { return fabsf(a); }
RKDG_HD inline double absR(double a) { return fabs(a); }

RKDG_HD inline float rcpR(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
RKDG_HD inline double rcpR(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}
RKDG_HD inline float sqrtR(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}
RKDG_HD inline double sqrtR(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    // s ~ sqrt(x), h ~ 1/(2 sqrt(x)) from y = rsq(x); one coupled Newton step; then s + (x - s^2) h.
    // Positive normal inputs only (depths and speeds); 0 gives 0, a negative input NaN as sqrt does.
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double r = fma(-s, h, 0.5);
    s = fma(s, r, s), h = fma(h, r, h);
    const double d = fma(-s, s, x);
    return x > 0.0 ? fma(d, h, s) : (x == 0.0 ? 0.0 : __builtin_nan(""));
#else
    return std::sqrt(x);
#endif
}
RKDG_HD inline float cbrt_rcpR(float x) {  // x^(-1/3)
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_exp2f(-0.333333343f * __builtin_amdgcn_logf(x));
#else
    return std::pow(x, -1.0f / 3.0f);
#endif
}
RKDG_HD inline double cbrt_rcpR(double x) {  // x^(-1/3): y <- y (4 - x y^3) / 3, quadratic
    double y = (double)cbrt_rcpR((float)x);
    y = y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
    return y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
}

// SPH: the spherical correction sp enters as in llf (wave speed sqrt(g h ((nx sp)^2 + ny^2)), the
// in-side flux's x-terms scaled); otherwise sp = 1 and nx^2 + ny^2 = 1.
template <typename Real, bool SPH = false>
RKDG_HD inline void llf_fast(Real grav, const Real* qi, const Real* qx, Real bath, Real nx, Real ny, Real* F,
                             Real sp = Real(1)) {
    const Real h_in = qi[0] + bath, h_ex = qx[0] + bath;
    const Real r_in = rcpR(depth_for_div(h_in)), r_ex = rcpR(depth_for_div(h_ex));
    const Real u_in = qi[1] * r_in, v_in = qi[2] * r_in;
    const Real u_ex = qx[1] * r_ex, v_ex = qx[2] * r_ex;
    const Real un_in = u_in * nx + v_in * ny, un_ex = u_ex * nx + v_ex * ny;
    const Real spc = SPH ? (nx * sp) * (nx * sp) + ny * ny : Real(1);
    const Real ci = grav * depth_for_speed(h_in), cx = grav * depth_for_speed(h_ex);
    const Real lam = SPH ? rmax(absR(un_in) + sqrtR(ci * spc), absR(un_ex) + sqrtR(cx * spc))
                         : rmax(absR(un_in) + sqrtR(ci), absR(un_ex) + sqrtR(cx));
    const Real pe_in = grav * qi[0] * (Real(0.5) * qi[0] + bath), pe_ex = grav * qx[0] * (Real(0.5) * qx[0] + bath);
    const Real fi0 = (SPH ? sp * qi[1] : qi[1]) * nx + qi[2] * ny, fe0 = qx[1] * nx + qx[2] * ny;
    const Real fi1 = u_in * fi0 + (SPH ? sp * pe_in : pe_in) * nx, fe1 = u_ex * fe0 + pe_ex * nx;
    const Real fi2 = v_in * fi0 + pe_in * ny, fe2 = v_ex * fe0 + pe_ex * ny;
    F[0] = Real(0.5) * (fi0 + fe0 + lam * (qi[0] - qx[0]));
    F[1] = Real(0.5) * (fi1 + fe1 + lam * (qi[1] - qx[1]));
    F[2] = Real(0.5) * (fi2 + fe2 + lam * (qi[2] - qx[2]));
}

// The same LLF flux with the pressures of both states given relative to a hydrostatic
// reference r: pi = pe_in - r, px = pe_ex - r.  The mass component and lambda are the plain
// ones; the momentum components come out as F - r n.
template <typename Real, bool SPH = false>
RKDG_HD inline void llf_rel(Real grav, const Real* qi, const Real* qx, Real bath, Real pi, Real px, Real nx,
                            Real ny, Real* F, Real sp = Real(1)) {
    const Real h_in = qi[0] + bath, h_ex = qx[0] + bath;
    const Real r_in = rcpR(depth_for_div(h_in)), r_ex = rcpR(depth_for_div(h_ex));
    const Real u_in = qi[1] * r_in, v_in = qi[2] * r_in;
    const Real u_ex = qx[1] * r_ex, v_ex = qx[2] * r_ex;
    const Real un_in = u_in * nx + v_in * ny, un_ex = u_ex * nx + v_ex * ny;
    const Real spc = SPH ? (nx * sp) * (nx * sp) + ny * ny : Real(1);
    const Real ci = grav * depth_for_speed(h_in), cx = grav * depth_for_speed(h_ex);
    const Real lam = SPH ? rmax(absR(un_in) + sqrtR(ci * spc), absR(un_ex) + sqrtR(cx * spc))
                         : rmax(absR(un_in) + sqrtR(ci), absR(un_ex) + sqrtR(cx));
    const Real fi0 = (SPH ? sp * qi[1] : qi[1]) * nx + qi[2] * ny, fe0 = qx[1] * nx + qx[2] * ny;
    F[0] = Real(0.5) * (fi0 + fe0 + lam * (qi[0] - qx[0]));
    F[1] = Real(0.5) * (u_in * fi0 + u_ex * fe0 + ((SPH ? sp * pi : pi) + px) * nx + lam * (qi[1] - qx[1]));
    F[2] = Real(0.5) * (v_in * fi0 + v_ex * fe0 + (pi + px) * ny + lam * (qi[2] - qx[2]));
}

// Interface flux of one side in the relative-pressure form: pi/px as for llf_rel, bl the
// bathymetry the LLF flux uses (the in side's), bi/bx each side's own (for the wet/dry fixes), r
// the caller's hydrostatic reference at this point.  Returns the flux of the side given first
// (want_in) or second, w.r.t. that side's outward normal, with momentum relative to r.  The
// wet/dry fixes of rkdg_swe_is_internal.hpp:47-90 (rare) are applied to the plain flux, rebuilt as
// F_rel + r n; they depend only on the sign of the mass flux and on which side is dry, so they
// give each side the same flux whichever side is passed first.
template <typename Real, bool WD, bool SPH = false>
RKDG_HD inline void interface_flux_fast(Real grav, const Real* qi, const Real* qx, Real pi, Real px, Real bl, Real bi,
                                        Real bx, Real nx, Real ny, bool wet_in, bool wet_ex, bool want_in, Real r,
                                        Real* F, Real sp = Real(1)) {
    Real Fi[3];
    llf_rel<Real, SPH>(grav, qi, qx, bl, pi, px, nx, ny, Fi, sp);
    if (WD && ((Fi[0] > Real(1e-12) && !(wet_in && wet_ex)) || (Fi[0] < -Real(1e-12) && !(wet_in && wet_ex)))) {
        Fi[1] += r * nx, Fi[2] += r * ny;
        Real Fe[3] = {-Fi[0], -Fi[1], -Fi[2]};
        if (Fi[0] > Real(1e-12)) {
            if (!wet_in) {
                Fi[0] = 0, Fi[1] = 0, Fi[2] = 0;
                const Real tx = ny, ty = -nx, mx = -nx, my = -ny;
                const Real qn_ex = -(qx[1] * mx + qx[2] * my), qt_ex = qx[1] * tx + qx[2] * ty;
                const Real qe[3] = {qx[0], qn_ex * mx + qt_ex * tx, qn_ex * my + qt_ex * ty};
                llf_fast<Real, SPH>(grav, qx, qe, bx, mx, my, Fe, sp);
            } else {  // !wet_ex
                const Real f0 = Fi[0];
                llf_fast<Real, SPH>(Real(0.0), qx, qi, bx, -nx, -ny, Fe, sp);
                Fe[0] = -f0;
            }
        } else {
            if (!wet_ex) {
                Fe[0] = 0, Fe[1] = 0, Fe[2] = 0;
                const Real tx = -ny, ty = nx;
                const Real qn_ex = -(qi[1] * nx + qi[2] * ny), qt_ex = qi[1] * tx + qi[2] * ty;
                const Real qe[3] = {qi[0], qn_ex * nx + qt_ex * tx, qn_ex * ny + qt_ex * ty};
                llf_fast<Real, SPH>(grav, qi, qe, bi, nx, ny, Fi, sp);
            } else {  // !wet_in
                const Real fe0 = Fe[0];
                llf_fast<Real, SPH>(Real(0.0), qi, qx, bi, nx, ny, Fi, sp);
                Fi[0] = -fe0;
            }
        }
        const Real s = want_in ? Real(1.0) : -Real(1.0);
        const Real* src = want_in ? Fi : Fe;
        F[0] = src[0], F[1] = src[1] - r * s * nx, F[2] = src[2] - r * s * ny;
        return;
    }
    const Real s = want_in ? Real(1.0) : -Real(1.0);
    F[0] = s * Fi[0], F[1] = s * Fi[1], F[2] = s * Fi[2];
}

// Exterior state of a tide, flow or function boundary point.  Tide/flow: rkdg_swe_bc_tide.hpp:70-96, rkdg_swe_bc_flow.hpp:72-101;
// function: rkdg_swe_bc_function.hpp:139-172.
template <typename Real, int P, bool CF>
RKDG_HD inline void exterior_state(const StageArgs<Real>& A, int bc, int e, int ed, int f, int g, const Real* qi,
                                   Real nx, Real ny, Real* qe, double t, double ramp) {
    constexpr int NE = Sizes<P>::NE;
    if (bc == 1 || bc == 2) {
        const Real acc = harmonic_forcing<Real, NE>(A, ed, g, t, ramp);
        if (bc == 1)
            qe[0] = acc, qe[1] = qi[1], qe[2] = qi[2];
        else
            qe[0] = qi[0], qe[1] = acc * nx, qe[2] = acc * ny;
    } else if (CF) {
        using C = ConstTables<Real, P>;
        const int ne = A.ne;
        const Real x = A.vxy[e] * C::psie[f * 3][g] + A.vxy[(size_t)ne + e] * C::psie[f * 3 + 1][g] +
                       A.vxy[2 * (size_t)ne + e] * C::psie[f * 3 + 2][g];
        const Real y = A.vxy[3 * (size_t)ne + e] * C::psie[f * 3][g] + A.vxy[4 * (size_t)ne + e] * C::psie[f * 3 + 1][g] +
                       A.vxy[5 * (size_t)ne + e] * C::psie[f * 3 + 2][g];
        problem_ic<Real>(A.problem, A.pparam, A.g, t, x, y, qe);
    } else {
        qe[0] = qi[0], qe[1] = qi[1], qe[2] = qi[2];  // unreachable: function BCs launch the CF build
    }
}

// r + a * c for a compile-time table entry c; entries that are exactly zero (derivatives of the
// constant mode, basis functions vanishing at a point) fold away at compile time.
template <typename Real>
RKDG_HD inline Real fmac_tab(Real a, Real c, Real r) { return c != Real(0) ? fmaR(a, c, r) : r; }

template <typename Real, int P, bool WD, bool CF, bool ENS>
RKDG_HD RKDG_FORCE_INLINE inline void stage_element_fast(const StageArgs<Real>& A, const Real* phie_lds, const Real* phiew_lds,
                                       int e, const InstScal<Real>& IS, const Real* t0_lds) {
    using C = ConstTables<Real, P>;
    constexpr int ND = Sizes<P>::ND, NG = Sizes<P>::NG, NE = Sizes<P>::NE;
    const int ne = A.ne;
    const Real grav = A.g;

    // The bathymetry is the L2 projection of the nodal depths' linear interpolant
    // (swe_pre_init_data.hpp:20-24): exactly linear, so its modes above the three linear ones are
    // rounding noise.  The fast forms use the linear modes only, in the volume and on every face
    // alike (the hydrostatic balance below needs one representation of b everywhere); its
    // gradient is then one constant per element.
    constexpr int NB = 3;

    // ---- every load this element needs, issued before any arithmetic
    Real q[3][ND], bath[ND];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) q[k][d] = ldc(A.q_s, ne, k * ND + d, e);
#pragma unroll
    for (int d = 0; d < ND; ++d) bath[d] = d < NB ? ldc(A.bath, ne, d, e) : Real(0);
    const Real J00 = ldc(A.geom, ne, 0, e), J01 = ldc(A.geom, ne, 1, e), J10 = ldc(A.geom, ne, 2, e),
                J11 = ldc(A.geom, ne, 3, e), adet = ldc(A.geom, ne, 4, e);
    const uint32_t meta = ldc(A.meta, ne, 0, e);
    int nbv[3], ed[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) nbv[f] = ldc(A.nb, ne, f, e), ed[f] = ldc(A.edge, ne, f, e);
    // The first earlier state of the SSP-RK combination (q0, read by the update at the end) is
    // fetched into LDS (t0_lds) when the stage starts, so it holds no registers while in flight.
    // Neighbour modes and edge geometry are loaded at the face itself, which keeps one face's
    // worth of neighbour registers live: FP32 113 VGPRs (4 waves per SIMD; 155 and 3 waves when
    // it held q0 and the next face's modes in registers), FP64 215 (2 waves).
    auto load_nb = [&](int f, Real (&qm)[3][ND], Real (&bm)[ND]) {
        const int n = nbv[f] >= 0 ? nbv[f] : e;  // boundary faces read their own (cached) modes
#pragma unroll
        for (int d = 0; d < ND; ++d) {
#pragma unroll
            for (int k = 0; k < 3; ++k) qm[k][d] = ldc(A.q_s, ne, k * ND + d, n);
            bm[d] = d < NB ? ldc(A.bath, ne, d, n) : Real(0);
        }
    };
    Real qn_m[3][ND], bn_m[ND];
    const bool wet = WD ? (A.wet_in[e] != 0) : true;
    bool wet_nb[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) wet_nb[f] = (WD && nbv[f] >= 0) ? (A.wet_in[nbv[f]] != 0) : true;
    const Real gmn2 = A.friction == 2 ? ldc(A.gmn2, ne, 0, e) : Real(0.0);
    const Real rdet = rcpR(adet);

    // ---- hydrostatic reference.  A wet element works with its momentum pressure relative to
    // g zr (zr/2 + b), zr = its mean surface elevation.  With the quadratures exact for these
    // degrees, volume + source of that part equal a face integral of g zr (zr/2 + b_own) n, so
    // the volume uses pe - g zr (zr/2 + b) = g z' (h - z'/2) with z' = ze - zr, the source
    // g z' grad b, and every face subtracts g zr (zr/2 + b_own) n.  Same discretisation; the
    // large canceling terms (g ze b with b up to kilometres) never form in FP32.
    constexpr Real PHI0 = C::phi[0][0];  // the constant mode's value
    const Real zc = q[0][0] * PHI0;      // mean surface elevation (mode 0)
    // With the spherical correction the volume x-fluxes carry sp and the source does not, so the
    // hydrostatic part no longer cancels against a face integral: spherical runs keep the plain
    // pressures (reference 0), as dry elements do.
    const bool sph = CF && A.sp_vol != nullptr;
    const Real zr = (wet && !sph) ? zc : Real(0.0);
    const Real zv = sph ? zc : Real(0.0);  // added to the volume's elevation relative to zc

    // ---- volume and source: rhs (unscaled by |det J|) = sum_g [idphi . (J^-T F) + ipf . S]
    Real rhs[3][ND];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) rhs[k][d] = Real(0.0);
    if (wet) {
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            Real zp = Real(0.0), q1 = Real(0.0), q2 = Real(0.0), bg = Real(0.0), bz1 = Real(0.0), bz2 = Real(0.0);
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                if (d > 0) zp = fmac_tab(q[0][d], C::phi[d][g], zp);
                q1 = fmac_tab(q[1][d], C::phi[d][g], q1);
                q2 = fmac_tab(q[2][d], C::phi[d][g], q2);
                if (d < NB) {
                    bg = fmac_tab(bath[d], C::phi[d][g], bg);
                    bz1 = fmac_tab(bath[d], C::dphi0[d][g], bz1);
                    bz2 = fmac_tab(bath[d], C::dphi1[d][g], bz2);
                }
            }
            const Real h = (zc + bg) + zp, rh = rcpR(h);
            const Real u = q1 * rh, v = q2 * rh;
            const Real zpv = zv + zp;  // elevation relative to the volume's reference (zc, or 0)
            const Real pe = grav * zpv * (h - Real(0.5) * zpv);  // relative to g zc (zc/2 + b), or plain
            const Real uvh = u * q2;
            Real fx[3] = {q1, u * q1 + pe, uvh};
            const Real fy[3] = {q2, uvh, v * q2 + pe};
            if (sph) {  // Fx *= sp (rkdg_swe_proc_volume.hpp:25-30)
                const Real sp = ldc(A.sp_vol, ne, g, e);
                fx[0] *= sp, fx[1] *= sp, fx[2] *= sp;
            }
            Real G1[3], G2[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) G1[k] = J00 * fx[k] + J01 * fy[k], G2[k] = J10 * fx[k] + J11 * fy[k];
            const Real gz = grav * (zv + zp);
            Real S[3] = {Real(0.0), gz * (bz1 * J00 + bz2 * J10), gz * (bz1 * J01 + bz2 * J11)};
            if (CF && A.fsrc) {
                const Real x = A.vxy[e] * C::psi[0][g] + A.vxy[(size_t)ne + e] * C::psi[1][g] +
                                A.vxy[2 * (size_t)ne + e] * C::psi[2][g];
                const Real y = A.vxy[3 * (size_t)ne + e] * C::psi[0][g] + A.vxy[4 * (size_t)ne + e] * C::psi[1][g] +
                                A.vxy[5 * (size_t)ne + e] * C::psi[2][g];
                Real fsv[3];
                problem_source<Real>(IS.t, x, y, fsv);
                S[0] += fsv[0], S[1] += fsv[1], S[2] += fsv[2];
            }
            if (A.friction) {
                Real Cf = A.cf;
                if (A.friction == 2) {
                    Cf = gmn2 * cbrt_rcpR(h);
                    if (Cf < A.cf) Cf = A.cf;
                }
                const Real tau = Cf * sqrtR(u * u + v * v) * rh;
                S[1] -= tau * q1;
                S[2] -= tau * q2;
            }
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Real r = rhs[k][d];
                    r = fmac_tab(G1[k], C::idphi0[g][d], r);
                    r = fmac_tab(G2[k], C::idphi1[g][d], r);
                    if (CF || k > 0) r = fmac_tab(S[k], C::ipf[g][d], r);  // S[0] = 0 without CF
                    rhs[k][d] = r;
                }
        }
    }

    // ---- interior faces (local face order; own tables compile-time, the neighbour's from LDS).
    // Each side evaluates the LLF flux from its own side: its own trace first, its outward normal,
    // the in side's bathymetry.  Both sides form the surface trace as zc + (sum of modes 1..) in
    // the same order, so the two evaluations see the same traces, normals that are exact
    // negations and the same lambda; the mass flux, a sum of two normal fluxes and lambda times a
    // trace difference, comes out exactly negated on the other side (mass is conserved to
    // roundoff), the momentum components agree to rounding.  Builds with the closed-form problem
    // functions (CF) carry the spherical correction, which scales the in side's normal flux only:
    // they keep the in/ex roles (the lower element ID evaluates as the in side).
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        load_nb(f, qn_m, bn_m);
        if (nbv[f] >= 0 && !((meta >> (12 + f)) & 1) && (wet || wet_nb[f])) {
            const Real fnx = ldc(A.egeom, A.nedge, 0, ed[f]);
            const Real fny = ldc(A.egeom, A.nedge, 1, ed[f]);
            const Real fsJ = ldc(A.egeom, A.nedge, 2, ed[f]);
            const int fn = (meta >> (4 * f + 2)) & 3;
            const Real* pn = phie_lds + fn * ND * NE;
            const bool is_in = e < nbv[f];
            const Real zcn = qn_m[0][0] * PHI0;
            const Real dzr = zcn - zr;
            const Real mx = is_in ? fnx : -fnx, my = is_in ? fny : -fny;  // this element's outward normal
            Real F[3][NE];
#pragma unroll
            for (int g = 0; g < NE; ++g) {
                const int gm = NE - 1 - g;
                Real zo = Real(0.0), o1 = Real(0.0), o2 = Real(0.0), bo = Real(0.0), zn = Real(0.0), n1 = Real(0.0), n2 = Real(0.0), bn = Real(0.0);
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const Real w = C::phie[f * ND + d][g];
                    const Real wn = pn[d * NE + gm];
                    if (d > 0) zo = fmac_tab(q[0][d], w, zo), zn = fmaR(qn_m[0][d], wn, zn);
                    o1 = fmac_tab(q[1][d], w, o1), o2 = fmac_tab(q[2][d], w, o2);
                    if (d < NB) bo = fmac_tab(bath[d], w, bo);
                    n1 = fmaR(qn_m[1][d], wn, n1), n2 = fmaR(qn_m[2][d], wn, n2);
                    if (d < NB) bn = fmaR(bn_m[d], wn, bn);
                }
                const Real own[3] = {zc + zo, o1, o2}, nbs[3] = {zcn + zn, n1, n2};
                // surface elevations relative to zr, and the pressures relative to g zr (zr/2 + bo)
                const Real zpo = (wet && !sph) ? zo : own[0];
                const Real zpn = dzr + zn;
                const Real B = is_in ? bo : bn;        // the LLF uses the in-side bathymetry
                const Real D = is_in ? Real(0.0) : zr * (bn - bo);
                const Real po = grav * (zpo * (zr + Real(0.5) * zpo + B) + D);
                const Real pnb = grav * (zpn * (zr + Real(0.5) * zpn + B) + D);
                const Real rref = grav * zr * (Real(0.5) * zr + bo);
                const Real sp = sph ? ldc(A.sp_edge, A.nedge, is_in ? g : gm, ed[f]) : Real(1);
                Real Fg[3];
                if (!CF)  // from this element's side: no per-lane exchange of the two sides' operands
                    interface_flux_fast<Real, WD, CF>(grav, own, nbs, po, pnb, B, bo, bn, mx, my, wet, wet_nb[f], true, rref,
                                                      Fg);
                else if (is_in)
                    interface_flux_fast<Real, WD, CF>(grav, own, nbs, po, pnb, bo, bo, bn, fnx, fny, wet, wet_nb[f], true, rref,
                                                      Fg, sp);
                else
                    interface_flux_fast<Real, WD, CF>(grav, nbs, own, pnb, po, bn, bn, bo, fnx, fny, wet_nb[f], wet, false, rref,
                                                      Fg, sp);
                F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
            }
            const Real scale = fsJ * rdet;
#pragma unroll
            for (int g = 0; g < NE; ++g) F[0][g] *= scale, F[1][g] *= scale, F[2][g] *= scale;
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Real r = rhs[k][d];
#pragma unroll
                    for (int g = 0; g < NE; ++g) r = fmac_tab(F[k][g], -C::phiew[f * ND + d][g], r);
                    rhs[k][d] = r;
                }
        }
    }

    // ---- levee faces (internal barriers; rare: rolled loop, tables from LDS).  The weir state is
    // evaluated on the plain traces and the hydrostatic reference subtracted afterwards.
    if (meta & (7u << 12)) {
#pragma unroll 1
        for (int f = 0; f < 3; ++f) {
            if (!((meta >> (12 + f)) & 1)) continue;
            const int nbe = f == 0 ? nbv[0] : (f == 1 ? nbv[1] : nbv[2]);
            const bool wnb = f == 0 ? wet_nb[0] : (f == 1 ? wet_nb[1] : wet_nb[2]);
            if (!(wet || wnb)) continue;
            const int fed = f == 0 ? ed[0] : (f == 1 ? ed[1] : ed[2]);
            const int fn = (meta >> (4 * f + 2)) & 3;
            const Real fnx = ldc(A.egeom, A.nedge, 0, fed), fny = ldc(A.egeom, A.nedge, 1, fed);
            const Real fsJ = ldc(A.egeom, A.nedge, 2, fed);
            const Real* pf = phie_lds + f * ND * NE;
            const Real* pn = phie_lds + fn * ND * NE;
            const Real* pw = phiew_lds + f * ND * NE;
            Real zn_m[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) zn_m[d] = ldc(A.q_s, ne, d, nbe);
            const bool is_in = e < nbe;
            const Real* lp = A.levee_gp + (size_t)A.eforce[fed] * 3 * NE;
            const Real onx = is_in ? fnx : -fnx, ony = is_in ? fny : -fny;
            Real F[3][NE];
#pragma unroll
            for (int g = 0; g < NE; ++g) {
                const int gm = NE - 1 - g, gi = is_in ? g : gm;
                Real zo = Real(0), o1 = Real(0), o2 = Real(0), bo = Real(0), zn = Real(0);
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const Real w = pf[d * NE + g];
                    if (d > 0) zo = fmaR(q[0][d], w, zo);
                    o1 = fmaR(q[1][d], w, o1), o2 = fmaR(q[2][d], w, o2);
                    if (d < NB) bo = fmaR(bath[d], w, bo);
                    zn = fmaR(zn_m[d], pn[d * NE + gm], zn);
                }
                const Real own[3] = {zc + zo, o1, o2}, nbs[3] = {zn, Real(0), Real(0)};
                Real Fg[3];
                levee_flux<Real>(grav, is_in, own, nbs, bo, fnx, fny, lp[gi], lp[NE + gi], lp[2 * NE + gi], wet, Fg,
                                 sph ? ldc(A.sp_edge, A.nedge, gi, fed) : Real(1));
                const Real rref = grav * zr * (Real(0.5) * zr + bo);
                F[0][g] = Fg[0], F[1][g] = Fg[1] - rref * onx, F[2][g] = Fg[2] - rref * ony;
            }
            const Real scale = fsJ * rdet;
#pragma unroll
            for (int g = 0; g < NE; ++g) F[0][g] *= scale, F[1][g] *= scale, F[2][g] *= scale;
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Real r = rhs[k][d];
#pragma unroll
                    for (int g = 0; g < NE; ++g) r = fmac_tab(F[k][g], -pw[d * NE + g], r);
                    rhs[k][d] = r;
                }
        }
    }

    // ---- boundary faces (only for wet elements; rolled loop, tables from LDS)
    if (wet && ((nbv[0] | nbv[1] | nbv[2]) < 0)) {
#pragma unroll 1
        for (int f = 0; f < 3; ++f) {
            if (nbv[f] >= 0) continue;
            const int bc = -1 - nbv[f];
            const Real* pf = phie_lds + f * ND * NE;
            const Real* pw = phiew_lds + f * ND * NE;
            const int fed = f == 0 ? ed[0] : (f == 1 ? ed[1] : ed[2]);
            const Real fnx = ldc(A.egeom, A.nedge, 0, fed);
            const Real fny = ldc(A.egeom, A.nedge, 1, fed);
            const Real fsJ = ldc(A.egeom, A.nedge, 2, fed);
            Real F[3][NE];
#pragma unroll
            for (int g = 0; g < NE; ++g) {
                Real zo = Real(0.0), o1 = Real(0.0), o2 = Real(0.0), bo = Real(0.0);
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const Real w = pf[d * NE + g];
                    if (d > 0) zo = fmaR(q[0][d], w, zo);
                    o1 = fmaR(q[1][d], w, o1), o2 = fmaR(q[2][d], w, o2);
                    if (d < NB) bo = fmaR(bath[d], w, bo);
                }
                const Real qi[3] = {zc + zo, o1, o2};
                // elevations relative to the reference zb: zc, or 0 (plain pressures) when spherical
                const Real zb = sph ? Real(0.0) : zc, zob = sph ? qi[0] : zo;
                Real qe[3], zpe = zob;  // exterior elevation (land, flow: the interior one)
                if (bc == 0) {
                    const Real tx = -fny, ty = fnx;
                    const Real qn_ex = -(qi[1] * fnx + qi[2] * fny), qt_ex = qi[1] * tx + qi[2] * ty;
                    qe[0] = qi[0], qe[1] = qn_ex * fnx + qt_ex * tx, qe[2] = qn_ex * fny + qt_ex * ty;
                } else {
                    exterior_state<Real, P, CF>(A, bc, e, fed, f, g, qi, fnx, fny, qe, IS.t, IS.ramp);
                    if (bc != 2) zpe = qe[0] - zb;  // tide, function: the exterior elevation is given
                }
                const Real po = grav * zob * (zb + Real(0.5) * zob + bo);
                const Real pe = grav * zpe * (zb + Real(0.5) * zpe + bo);
                Real Fg[3];
                llf_rel<Real, CF>(grav, qi, qe, bo, po, pe, fnx, fny, Fg, sph ? ldc(A.sp_edge, A.nedge, g, fed) : Real(1));
                F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
            }
            const Real scale = fsJ * rdet;
#pragma unroll
            for (int g = 0; g < NE; ++g) F[0][g] *= scale, F[1][g] *= scale, F[2][g] *= scale;
#pragma unroll
            for (int d = 0; d < ND; ++d)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    Real r = rhs[k][d];
#pragma unroll
                    for (int g = 0; g < NE; ++g) r = fmac_tab(F[k][g], -pw[d * NE + g], r);
                    rhs[k][d] = r;
                }
        }
    }

    if (t0_lds) lds_columns_wait();
    // the first term's scalars are read once, outside the mode loop (see stage_element)
    const int nterm = A.nterm;
    const Real ta0 = A.term_a[0], tc0 = term_coef<ENS>(A, IS, 0), ca = A.cur_a, cc = cur_coef<ENS>(A, IS);
    const Real* const tq0 = A.term_q[0];
    const Real* const ts0 = A.term_sol[0];
    Real* const so = A.sol_out;
    Real nq[3][ND];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const Real sol = rhs[k][d] * C::minv[0][d];
            if (so) stc(so, ne, k * ND + d, e, sol);
            Real acc = ca * q[k][d] + cc * sol;
            if (nterm > 0) {
                Real term = ta0 * (t0_lds ? t0_lds[64 * (k * ND + d) + lane_index()] : ldc(tq0, ne, k * ND + d, e));
                if (ts0) term += tc0 * ldc(ts0, ne, k * ND + d, e);
                acc += term;
                for (int i = 1; i < nterm; ++i) {
                    term = A.term_a[i] * ldc(A.term_q[i], ne, k * ND + d, e);
                    if (A.term_sol[i]) term += term_coef<ENS>(A, IS, i) * ldc(A.term_sol[i], ne, k * ND + d, e);
                    acc += term;
                }
            }
            nq[k][d] = acc;
        }
    finish_element<Real, P, WD>(A, e, wet, nq);
}

}  // namespace rkdg
