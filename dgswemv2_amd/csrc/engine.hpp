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

#include <cmath>

namespace rkdg {

constexpr double kPI = 3.14159265359;  // general_definitions.hpp:227

template <int P>
struct Sizes {
    static constexpr int ND = (P + 1) * (P + 2) / 2;            // modes
    static constexpr int NG = P == 1 ? 3 : (P == 2 ? 6 : 12);   // Dunavant rule 2P points
    static constexpr int NE = P + 1;                            // Gauss-Legendre rule 2P+1 points
};

// Reference-element tables in the arithmetic type of the kernel (master_triangle.tpp:5-87).
template <typename Real, int P>
struct Tables {
    static constexpr int ND = Sizes<P>::ND, NG = Sizes<P>::NG, NE = Sizes<P>::NE;
    Real phi[ND][NG];
    Real dphi[2][ND][NG];
    Real idphi[2][NG][ND];
    Real ipf[NG][ND];
    Real minv[ND];
    Real phie[3][ND][NE];
    Real we[NE];
    Real psi[3][NG];        // shape functions at the rule points (gp global coordinates)
    Real psie[3][3][NE];    // [face][node][gp] shape functions at the edge points
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
    const int* eforce;       // [nedge] forcing record or -1
    const Real* famp;        // [nforced][ncon][NE]
    const Real* fphase;      // [nforced][ncon][NE]
    int ncon;
    Real freq[4], ffact[4], eqarg[4];
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
    unsigned long long stage_key;  // step * nstages + stage of this launch
};

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
// Local Lax-Friedrichs flux (numerical_fluxes/rkdg_swe_LLF_flux.hpp:7-60) with sp = 1.
// bath is the bathymetry of the side whose aux the caller passes.

template <typename Real>
RKDG_HD inline Real rmax(Real a, Real b) {  // std::max semantics: (a < b) ? b : a
    return (a < b) ? b : a;
}
template <typename Real>
RKDG_HD inline Real rmin(Real a, Real b) {  // std::min semantics: (b < a) ? b : a
    return (b < a) ? b : a;
}

template <typename Real>
RKDG_HD inline void llf(Real grav, const Real* qi, const Real* qx, Real bath, Real nx, Real ny, Real* F) {
    using std::fabs;
    using std::sqrt;
    const Real h_in = qi[0] + bath;
    const Real u_in = qi[1] / h_in;
    const Real v_in = qi[2] / h_in;
    const Real h_ex = qx[0] + bath;
    const Real u_ex = qx[1] / h_ex;
    const Real v_ex = qx[2] / h_ex;
    const Real un_in = u_in * nx + v_in * ny;
    const Real un_ex = u_ex * nx + v_ex * ny;
    const Real spc = nx * nx + ny * ny;
    const Real lam = rmax(fabs(un_in) + sqrt(grav * h_in * spc), fabs(un_ex) + sqrt(grav * h_ex * spc));
    const Real uuh_in = u_in * qi[1], vvh_in = v_in * qi[2], uvh_in = u_in * qi[2];
    const Real pe_in = grav * ((qi[0] * qi[0]) / 2 + qi[0] * bath);
    const Real uuh_ex = u_ex * qx[1], vvh_ex = v_ex * qx[2], uvh_ex = u_ex * qx[2];
    const Real pe_ex = grav * ((qx[0] * qx[0]) / 2 + qx[0] * bath);
    const Real fi0 = qi[1] * nx + qi[2] * ny;
    const Real fi1 = (uuh_in + pe_in) * nx + uvh_in * ny;
    const Real fi2 = uvh_in * nx + (vvh_in + pe_in) * ny;
    const Real fe0 = qx[1] * nx + qx[2] * ny;
    const Real fe1 = (uuh_ex + pe_ex) * nx + uvh_ex * ny;
    const Real fe2 = uvh_ex * nx + (vvh_ex + pe_ex) * ny;
    F[0] = Real(0.5) * (fi0 + fe0 + lam * (qi[0] - qx[0]));
    F[1] = Real(0.5) * (fi1 + fe1 + lam * (qi[1] - qx[1]));
    F[2] = Real(0.5) * (fi2 + fe2 + lam * (qi[2] - qx[2]));
}

// Reflective (land) exterior state + LLF (boundary_conditions/rkdg_swe_bc_land.hpp:68-86).
template <typename Real>
RKDG_HD inline void land_llf(Real grav, Real nx, Real ny, const Real* qi, Real bath, Real* F) {
    const Real tx = -ny, ty = nx;
    const Real qn_ex = -(qi[1] * nx + qi[2] * ny);
    const Real qt_ex = qi[1] * tx + qi[2] * ty;
    Real qe[3] = {qi[0], qn_ex * nx + qt_ex * tx, qn_ex * ny + qt_ex * ty};
    llf(grav, qi, qe, bath, nx, ny, F);
}

// Interface flux seen by one side (ISP::Internal::ComputeFlux, rkdg_swe_is_internal.hpp:24-94).
// qi/qx: in- and ex-side traces at one in-side point and its mirrored ex-side point;
// bi/bx: their bathymetry.  Returns the flux of the requested side.
template <typename Real, bool WD>
RKDG_HD inline void interface_flux(Real grav, const Real* qi, const Real* qx, Real bi, Real bx, Real nx, Real ny,
                                   bool wet_in, bool wet_ex, bool want_in, Real* F) {
    Real Fi[3], Fe[3];
    llf(grav, qi, qx, bi, nx, ny, Fi);
    Fe[0] = -Fi[0], Fe[1] = -Fi[1], Fe[2] = -Fi[2];
    if (WD) {
        if (Fi[0] > Real(1e-12)) {
            if (!wet_in) {
                Fi[0] = 0, Fi[1] = 0, Fi[2] = 0;
                land_llf(grav, -nx, -ny, qx, bx, Fe);
            } else if (!wet_ex) {
                const Real f0 = Fi[0];
                llf(Real(0), qx, qi, bx, -nx, -ny, Fe);
                Fe[0] = -f0;
            }
        } else if (Fi[0] < Real(-1e-12)) {
            if (!wet_ex) {
                Fe[0] = 0, Fe[1] = 0, Fe[2] = 0;
                land_llf(grav, nx, ny, qi, bi, Fi);
            } else if (!wet_in) {
                const Real fe0 = Fe[0];
                llf(Real(0), qi, qx, bi, nx, ny, Fi);
                Fi[0] = -fe0;
            }
        }
    }
    const Real* src = want_in ? Fi : Fe;
    F[0] = src[0], F[1] = src[1], F[2] = src[2];
}

// ---------------------------------------------------------------------------------------
// Wetting-drying of one element on its linear (vertex) state, in FP64
// (problem_postprocessor/swe_post_wet_dry.hpp:5-145).  qv[var][vertex] in/out; returns the
// new wet flag; *rewrite says whether the modal state must be rebuilt from qv.
RKDG_HD inline bool wet_dry_vertices(double qv[3][3], const double* bv, double bmin, bool wet, double h_o,
                                     double thr, bool* rewrite) {
    *rewrite = false;
    double h[3];
    unsigned n_dry = 0;
    for (int v = 0; v < 3; ++v) {
        h[v] = qv[0][v] + bv[v];
        if (h[v] <= h_o + thr) ++n_dry;
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
        if (h_avg <= h_o + thr) {
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
                if (h[v] <= h_o + thr) {
                    nd++;
                    dqx += qv[1][v], dqy += qv[2][v];
                    qv[1][v] = 0.0, qv[2][v] = 0.0;
                }
            for (int v = 0; v < 3; ++v) {
                qv[0][v] = h[v] - bv[v];
                if (h[v] > h_o + thr) {
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
        if (qv[0][hmax] > h_o - bmin) {
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
// The fused stage of one element.

template <typename Real, int P, bool WD>
RKDG_HD inline void stage_element(const StageArgs<Real>& A, const Tables<Real, P>& T, int e) {
    constexpr int ND = Sizes<P>::ND, NG = Sizes<P>::NG, NE = Sizes<P>::NE;
    const int ne = A.ne;
    const bool wet = WD ? (A.wet_in[e] != 0) : true;
    const Real grav = A.g;

    Real q[3][ND];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < ND; ++d) q[k][d] = A.q_s[(size_t)(k * ND + d) * ne + e];
    Real bath[ND];
    for (int d = 0; d < ND; ++d) bath[d] = A.bath[(size_t)d * ne + e];
    const Real J00 = A.geom[e], J01 = A.geom[(size_t)ne + e], J10 = A.geom[2 * (size_t)ne + e],
               J11 = A.geom[3 * (size_t)ne + e], adet = A.geom[4 * (size_t)ne + e];

    Real rhs[3][ND];
    if (wet) {
        // ---- volume_kernel (rkdg_swe_proc_volume.hpp:8-35): rhs = Fx*Dx + Fy*Dy
        Real ax[3][ND], ay[3][ND];
        for (int g = 0; g < NG; ++g) {
            Real qg[3];
            for (int k = 0; k < 3; ++k) {
                Real s = q[k][0] * T.phi[0][g];
                for (int d = 1; d < ND; ++d) s = s + q[k][d] * T.phi[d][g];
                qg[k] = s;
            }
            Real bg = bath[0] * T.phi[0][g];
            for (int d = 1; d < ND; ++d) bg = bg + bath[d] * T.phi[d][g];
            const Real h = qg[0] + bg;
            const Real u = qg[1] / h, v = qg[2] / h;
            const Real uuh = u * qg[1], vvh = v * qg[2], uvh = u * qg[2];
            const Real pe = grav * (Real(0.5) * (qg[0] * qg[0]) + qg[0] * bg);
            const Real fx[3] = {qg[1], uuh + pe, uvh};
            const Real fy[3] = {qg[2], uvh, vvh + pe};
            for (int d = 0; d < ND; ++d) {
                Real dx = Real(0.0);
                dx += T.idphi[0][g][d] * J00;
                dx += T.idphi[1][g][d] * J10;
                dx *= adet;
                Real dy = Real(0.0);
                dy += T.idphi[0][g][d] * J01;
                dy += T.idphi[1][g][d] * J11;
                dy *= adet;
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
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < ND; ++d) rhs[k][d] = ax[k][d] + ay[k][d];

        // ---- source_kernel (rkdg_swe_proc_source.hpp:8-18, swe_source.hpp:8-97)
        Real ts[3][ND];
        for (int g = 0; g < NG; ++g) {
            Real qg[3];
            for (int k = 0; k < 3; ++k) {
                Real s = q[k][0] * T.phi[0][g];
                for (int d = 1; d < ND; ++d) s = s + q[k][d] * T.phi[d][g];
                qg[k] = s;
            }
            Real dbx, dby;
            {
                Real sx = Real(0), sy = Real(0);
                for (int d = 0; d < ND; ++d) {
                    Real ex = Real(0.0);
                    ex += T.dphi[0][d][g] * J00;
                    ex += T.dphi[1][d][g] * J10;
                    Real ey = Real(0.0);
                    ey += T.dphi[0][d][g] * J01;
                    ey += T.dphi[1][d][g] * J11;
                    if (d == 0) {
                        sx = bath[0] * ex;
                        sy = bath[0] * ey;
                    } else {
                        sx = sx + bath[d] * ex;
                        sy = sy + bath[d] * ey;
                    }
                }
                dbx = sx, dby = sy;
            }
            Real src[3] = {Real(0.0), grav * (dbx * qg[0]), grav * (dby * qg[0])};
            if (A.fsrc) {
                const Real x = (A.vxy[e] * T.psi[0][g] + A.vxy[(size_t)ne + e] * T.psi[1][g]) +
                               A.vxy[2 * (size_t)ne + e] * T.psi[2][g];
                const Real y = (A.vxy[3 * (size_t)ne + e] * T.psi[0][g] + A.vxy[4 * (size_t)ne + e] * T.psi[1][g]) +
                               A.vxy[5 * (size_t)ne + e] * T.psi[2][g];
                Real f[3];
                problem_source<Real>(A.t, x, y, f);
                for (int k = 0; k < 3; ++k) src[k] = src[k] + f[k];
            }
            if (A.friction) {
                Real bg = bath[0] * T.phi[0][g];
                for (int d = 1; d < ND; ++d) bg = bg + bath[d] * T.phi[d][g];
                const Real h = qg[0] + bg;
                const Real u = qg[1] / h, v = qg[2] / h;
                Real Cf = A.cf;
                if (A.friction == 2) {
                    using std::pow;
                    Cf = A.gmn2[e] / pow(h, Real(1.0 / 3.0));
                    if (Cf < A.cf) Cf = A.cf;
                }
                using std::hypot;
                const Real tau = Cf * hypot(u, v) / h;
                src[1] -= tau * qg[1];
                src[2] -= tau * qg[2];
            }
            for (int d = 0; d < ND; ++d) {
                const Real w = T.ipf[g][d] * adet;
                for (int k = 0; k < 3; ++k) ts[k][d] = (g == 0) ? src[k] * w : ts[k][d] + src[k] * w;
            }
        }
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < ND; ++d) rhs[k][d] = rhs[k][d] + ts[k][d];
    } else {
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < ND; ++d) rhs[k][d] = Real(0.0);
    }

    // ---- faces: interface_kernel / boundary_kernel contributions, in reference order
    const uint32_t meta = A.meta[e];
    for (int s = 0; s < 3; ++s) {
        const int nbv = A.nb[(size_t)s * ne + e];
        const int ed = A.edge[(size_t)s * ne + e];
        const int f = (meta >> (4 * s)) & 3;
        const int fn = (meta >> (4 * s + 2)) & 3;
        const Real nx = A.egeom[ed], ny = A.egeom[(size_t)A.nedge + ed], sJ = A.egeom[2 * (size_t)A.nedge + ed];
        Real F[3][NE];
        if (nbv >= 0) {
            const bool wet_nb = WD ? (A.wet_in[nbv] != 0) : true;
            if (!(wet || wet_nb)) continue;
            const bool is_in = e < nbv;
            Real qo[3][NE], bo[NE];
            for (int g = 0; g < NE; ++g) {
                for (int k = 0; k < 3; ++k) {
                    Real acc = q[k][0] * T.phie[f][0][g];
                    for (int d = 1; d < ND; ++d) acc = acc + q[k][d] * T.phie[f][d][g];
                    qo[k][g] = acc;
                }
                Real acc = bath[0] * T.phie[f][0][g];
                for (int d = 1; d < ND; ++d) acc = acc + bath[d] * T.phie[f][d][g];
                bo[g] = acc;
            }
            Real qn[3][NE], bn[NE];
            {
                Real qnb[3][ND], bnb[ND];
                for (int k = 0; k < 3; ++k)
                    for (int d = 0; d < ND; ++d) qnb[k][d] = A.q_s[(size_t)(k * ND + d) * ne + nbv];
                for (int d = 0; d < ND; ++d) bnb[d] = A.bath[(size_t)d * ne + nbv];
                for (int g = 0; g < NE; ++g) {
                    for (int k = 0; k < 3; ++k) {
                        Real acc = qnb[k][0] * T.phie[fn][0][g];
                        for (int d = 1; d < ND; ++d) acc = acc + qnb[k][d] * T.phie[fn][d][g];
                        qn[k][g] = acc;
                    }
                    Real acc = bnb[0] * T.phie[fn][0][g];
                    for (int d = 1; d < ND; ++d) acc = acc + bnb[d] * T.phie[fn][d][g];
                    bn[g] = acc;
                }
            }
            for (int g = 0; g < NE; ++g) {
                const int gm = NE - 1 - g;  // mirrored point on the other side
                Real Fg[3];
                if (is_in) {
                    const Real qi[3] = {qo[0][g], qo[1][g], qo[2][g]};
                    const Real qx[3] = {qn[0][gm], qn[1][gm], qn[2][gm]};
                    interface_flux<Real, WD>(grav, qi, qx, bo[g], bn[gm], nx, ny, wet, wet_nb, true, Fg);
                } else {
                    const Real qi[3] = {qn[0][gm], qn[1][gm], qn[2][gm]};
                    const Real qx[3] = {qo[0][g], qo[1][g], qo[2][g]};
                    interface_flux<Real, WD>(grav, qi, qx, bn[gm], bo[g], nx, ny, wet_nb, wet, false, Fg);
                }
                F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
            }
        } else {
            if (!wet) continue;
            const int bc = -1 - nbv;
            Real qo[3][NE], bo[NE];
            for (int g = 0; g < NE; ++g) {
                for (int k = 0; k < 3; ++k) {
                    Real acc = q[k][0] * T.phie[f][0][g];
                    for (int d = 1; d < ND; ++d) acc = acc + q[k][d] * T.phie[f][d][g];
                    qo[k][g] = acc;
                }
                Real acc = bath[0] * T.phie[f][0][g];
                for (int d = 1; d < ND; ++d) acc = acc + bath[d] * T.phie[f][d][g];
                bo[g] = acc;
            }
            for (int g = 0; g < NE; ++g) {
                const Real qi[3] = {qo[0][g], qo[1][g], qo[2][g]};
                Real qe[3];
                if (bc == 0) {  // Land (rkdg_swe_bc_land.hpp:40-66)
                    const Real tx = -ny, ty = nx;
                    const Real qn_ex = -(qi[1] * nx + qi[2] * ny);
                    const Real qt_ex = qi[1] * tx + qi[2] * ty;
                    qe[0] = qi[0], qe[1] = qn_ex * nx + qt_ex * tx, qe[2] = qn_ex * ny + qt_ex * ty;
                } else if (bc == 1 || bc == 2) {  // Tide / Flow (rkdg_swe_bc_tide.hpp:70-96, rkdg_swe_bc_flow.hpp:72-101)
                    const int fr = A.eforce[ed];
                    Real acc = Real(0.0);
                    for (int c = 0; c < A.ncon; ++c) {
                        using std::cos;
                        const Real amp = A.famp[((size_t)fr * A.ncon + c) * NE + g];
                        const Real ph = A.fphase[((size_t)fr * A.ncon + c) * NE + g];
                        acc += Real(A.ramp) * A.ffact[c] * amp *
                               cos(A.freq[c] * Real(A.t) + (A.eqarg[c] - ph) * Real(kPI) / Real(180));
                    }
                    if (bc == 1)
                        qe[0] = acc, qe[1] = qi[1], qe[2] = qi[2];
                    else
                        qe[0] = qi[0], qe[1] = acc * nx, qe[2] = acc * ny;
                } else {  // Function (rkdg_swe_bc_function.hpp:139-172): ic_q(t, x_gp)
                    const Real x = (A.vxy[e] * T.psie[f][0][g] + A.vxy[(size_t)ne + e] * T.psie[f][1][g]) +
                                   A.vxy[2 * (size_t)ne + e] * T.psie[f][2][g];
                    const Real y =
                        (A.vxy[3 * (size_t)ne + e] * T.psie[f][0][g] + A.vxy[4 * (size_t)ne + e] * T.psie[f][1][g]) +
                        A.vxy[5 * (size_t)ne + e] * T.psie[f][2][g];
                    problem_ic<Real>(A.problem, A.pparam, grav, A.t, x, y, qe);
                }
                Real Fg[3];
                llf(grav, qi, qe, bo[g], nx, ny, Fg);
                F[0][g] = Fg[0], F[1][g] = Fg[1], F[2][g] = Fg[2];
            }
        }
        // rhs -= F * (phi_e * (w sJ))   (IntegrationPhiIN/EX, interface.hpp:170-182)
        for (int d = 0; d < ND; ++d) {
            Real acc[3];
            for (int g = 0; g < NE; ++g) {
                const Real w = T.phie[f][d][g] * (T.we[g] * sJ);
                for (int k = 0; k < 3; ++k) acc[k] = (g == 0) ? F[k][g] * w : acc[k] + F[k][g] * w;
            }
            for (int k = 0; k < 3; ++k) rhs[k][d] = rhs[k][d] - acc[k];
        }
    }

    // ---- M^-1 and the SSP-RK combination (element.hpp:500-505, explicit_ssp_rk_stepper.hpp:67-80)
    Real nq[3][ND];
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < ND; ++d) {
            const Real sol = rhs[k][d] * (T.minv[d] / adet);
            if (A.sol_out) A.sol_out[(size_t)(k * ND + d) * ne + e] = sol;
            Real acc = Real(0.0);
            for (int i = 0; i < A.nterm; ++i) {
                const size_t ix = (size_t)(k * ND + d) * ne + e;
                Real term = A.term_a[i] * A.term_q[i][ix];
                if (A.term_sol[i]) term = term + A.term_c[i] * A.term_sol[i][ix];
                acc = acc + term;
            }
            acc = acc + (A.cur_a * q[k][d] + A.cur_c * sol);
            nq[k][d] = acc;
        }

    // ---- wetting_drying_kernel on the new state (swe_post_wet_dry.hpp:5-145), FP64
    bool wet_new = wet;
    if (WD) {
        double qv[3][3];
        for (int k = 0; k < 3; ++k) {
            const double m[3] = {(double)nq[k][0], (double)nq[k][1], (double)nq[k][2]};
            to_vertices(m, qv[k]);
        }
        const double bv[3] = {A.bath_vrtx[e], A.bath_vrtx[(size_t)ne + e], A.bath_vrtx[2 * (size_t)ne + e]};
        const double bmin = rmin(rmin(bv[0], bv[1]), bv[2]);
        bool rewrite = false;
        wet_new = wet_dry_vertices(qv, bv, bmin, wet, A.h_o, A.h_thr, &rewrite);
        if (rewrite)
            for (int k = 0; k < 3; ++k) {
                double m[3];
                from_vertices(qv[k], m);
                for (int d = 0; d < 3; ++d) nq[k][d] = (Real)m[d];
                for (int d = 3; d < ND; ++d) nq[k][d] = Real(0.0);
            }
        A.wet_out[e] = wet_new ? 1 : 0;
    }

    bool has_nan = false;
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < ND; ++d) {
            A.q_out[(size_t)(k * ND + d) * ne + e] = nq[k][d];
            has_nan |= (nq[k][d] != nq[k][d]);
        }
    if (A.nan_check && has_nan) {
        const unsigned long long key = (A.stage_key << 32) | (unsigned long long)(unsigned)e;
#ifdef __HIP_DEVICE_COMPILE__
        atomicMin(A.nan_key, key);
#else
        if (key < *A.nan_key) *A.nan_key = key;
#endif
    }
}

}  // namespace rkdg
