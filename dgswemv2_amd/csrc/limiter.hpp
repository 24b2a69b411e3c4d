// Cockburn-Shu slope limiter and trouble check, per element (FP64 arithmetic in every form).
//
// The arithmetic is FP64 without fused multiply-adds in every form (each function turns contraction
// off, whatever the translation unit's setting), so a limiter decision depends only on the state.
//
// Three passes replace CS_slope_limiter_serial (problem_slope_limiter/swe_CS_sl_serial.hpp:9-23):
//   sl_prepare_element  - barycentre of the linear part (slope_limiting_prepare_element_kernel,
//                         swe_CS_slope_limiter.hpp:8-20) and the KXRCF-style indicator I of
//                         check_trouble (swe_trouble_check.hpp:13-104), written as the seed set;
//   sl_dilate_element   - one of the six neighbour-dilation passes (:106-115), double-buffered;
//   sl_limit_element    - slope_limiting_kernel (swe_CS_slope_limiter.hpp:88-189) reading the
//                         neighbours' barycentres from the prepare pass, then the NaN scan.
#pragma once

#include "engine.hpp"

namespace rkdg {

template <typename Real>
struct LimiterArgs {
    int ne, nedge;
    Real* q;                  // [3][ND][ne] state of the new stage, limited in place
    const Real* bath;         // [ND][ne]
    const int* nb;            // [3][ne] face slots (as in StageArgs)
    const int* edge;
    const uint32_t* meta;
    const Real* egeom;        // [3][nedge]
    const uint8_t* wet;       // [ne] flags after wetting-drying
    const double* bath_vrtx;  // [3][ne]
    const int* neighbor;      // [3][ne] by local face, -1 on a boundary
    const double* median;     // [6][ne]
    const double* alpha;      // [6][ne]
    const double* rsq;        // [3][ne]
    const double* len;        // [3][ne]
    const double* radius;     // [ne]
    const double* bath_bc;    // [ne]
    double* baryctr;          // [3][ne]
    uint8_t* set_a;           // [ne] troubled seed / dilation buffers
    uint8_t* set_b;
    uint8_t* troubled;        // [ne]
    uint8_t* limited;         // [3][ne]
    double g, M, nu;
    int p;
    int nan_check;
    unsigned long long* nan_key;
    unsigned long long stage_key;  // step * nstages + stage (the stage alone with step_ctr)
    const unsigned long long* step_ctr;  // device step counter of an ensemble, or null
    int key_ns;
};

template <typename Real, int P>
RKDG_HD inline void load_linear(const LimiterArgs<Real>& L, int e, double ql[3][3]) {
#pragma clang fp contract(off)
    constexpr int ND = Sizes<P>::ND;
    for (int k = 0; k < 3; ++k) {
        const double m[3] = {(double)L.q[(size_t)(k * ND) * L.ne + e], (double)L.q[(size_t)(k * ND + 1) * L.ne + e],
                             (double)L.q[(size_t)(k * ND + 2) * L.ne + e]};
        to_vertices(m, ql[k]);
    }
}

template <typename Real, int P>
RKDG_HD RKDG_FORCE_INLINE inline void sl_prepare_element(const LimiterArgs<Real>& L, const FaceTables<Real>& FT, int e) {
#pragma clang fp contract(off)
    constexpr int ND = Sizes<P>::ND, NE = Sizes<P>::NE;
    const int ne = L.ne;
    const bool wet = L.wet[e] != 0;
    double ql[3][3];
    load_linear<Real, P>(L, e, ql);
    if (wet)
        for (int k = 0; k < 3; ++k)
            L.baryctr[(size_t)k * ne + e] = (ql[k][0] * (1.0 / 3.0) + ql[k][1] * (1.0 / 3.0)) + ql[k][2] * (1.0 / 3.0);

    // check_trouble: interfaces of this element in (in, ex) order = face slots
    double I = 0.0, perimeter = 0.0;
    const uint32_t meta = L.meta[e];
    for (int s = 0; s < 3; ++s) {
        const int nbv = L.nb[(size_t)s * ne + e];
        if (nbv < 0) continue;
        const bool wet_nb = L.wet[nbv] != 0;
        if (!(wet || wet_nb)) continue;
        const int ed = L.edge[(size_t)s * ne + e];
        const int f = (meta >> (4 * s)) & 3, fn = (meta >> (4 * s + 2)) & 3;
        const bool is_in = e < nbv;
        const double nx = (double)L.egeom[ed], ny = (double)L.egeom[(size_t)L.nedge + ed];
        const double sJ = (double)L.egeom[2 * (size_t)L.nedge + ed];
        const double onx = is_in ? nx : -nx, ony = is_in ? ny : -ny;
        double ze_o[NE], qx_o[NE], h_o[NE], h_n[NE];
        for (int g = 0; g < NE; ++g) {
            double z = (double)L.q[(size_t)(0 * ND) * ne + e] * (double)FT.phie[((f) * ND + (0)) * NE + (g)];
            double x = (double)L.q[(size_t)(1 * ND) * ne + e] * (double)FT.phie[((f) * ND + (0)) * NE + (g)];
            double b = (double)L.bath[e] * (double)FT.phie[((f) * ND + (0)) * NE + (g)];
            double zn = (double)L.q[(size_t)(0 * ND) * ne + nbv] * (double)FT.phie[((fn) * ND + (0)) * NE + (g)];
            double bn = (double)L.bath[nbv] * (double)FT.phie[((fn) * ND + (0)) * NE + (g)];
            for (int d = 1; d < ND; ++d) {
                z = z + (double)L.q[(size_t)(0 * ND + d) * ne + e] * (double)FT.phie[((f) * ND + (d)) * NE + (g)];
                x = x + (double)L.q[(size_t)(1 * ND + d) * ne + e] * (double)FT.phie[((f) * ND + (d)) * NE + (g)];
                b = b + (double)L.bath[(size_t)d * ne + e] * (double)FT.phie[((f) * ND + (d)) * NE + (g)];
                zn = zn + (double)L.q[(size_t)(0 * ND + d) * ne + nbv] * (double)FT.phie[((fn) * ND + (d)) * NE + (g)];
                bn = bn + (double)L.bath[(size_t)d * ne + nbv] * (double)FT.phie[((fn) * ND + (d)) * NE + (g)];
            }
            ze_o[g] = z, qx_o[g] = x;
            h_o[g] = z + b;
            h_n[g] = zn + bn;
        }
        // quirk kept: q_n is formed from rows 0/1 of q (ze, qx) (swe_trouble_check.hpp:29-35)
        double integ = 0.0;
        for (int g = 0; g < NE; ++g) {
            const double qn = onx * ze_o[g] + ony * qx_o[g];
            const double w = (double)ConstTables<Real, P>::we[0][g] * sJ;
            integ = (g == 0) ? qn * w : integ + qn * w;
        }
        if (integ < 0.0) {
            double hi = 0.0;
            for (int g = 0; g < NE; ++g) {
                double hd = h_o[g];
                hd -= h_n[NE - 1 - g];
                const double w = (double)ConstTables<Real, P>::we[0][g] * sJ;
                hi = (g == 0) ? hd * w : hi + hd * w;
            }
            perimeter += L.len[(size_t)f * ne + e];
            I += (hi < 0 ? -hi : hi);
        }
    }
    uint8_t seed = 0;
    if (wet && perimeter != 0.0) {
        double h_norm = 0.0;
        for (int v = 0; v < 3; ++v) h_norm = rmax(h_norm, ql[0][v] + L.bath_vrtx[(size_t)v * ne + e]);
        using std::pow;
        I /= (pow(L.radius[e], (L.p + 1.0) / 2.0) * perimeter * h_norm);
        seed = I > 1.0 ? 1 : 0;
    }
    L.set_a[e] = seed;
    L.troubled[e] = 0;
}

RKDG_HD inline void sl_dilate_element(int ne, const int* neighbor, const uint8_t* in_set, uint8_t* out_set,
                                      uint8_t* troubled, int e) {
    if (in_set[e]) troubled[e] = 1;
    uint8_t any = 0;
    for (int k = 0; k < 3; ++k) {
        const int n = neighbor[(size_t)k * ne + e];
        if (n >= 0) any |= in_set[n];
    }
    out_set[e] = any;
}

RKDG_HD inline void mv3(const double A[3][3], const double* x, double* y) {
#pragma clang fp contract(off)
    for (int i = 0; i < 3; ++i) y[i] = (A[i][0] * x[0] + A[i][1] * x[1]) + A[i][2] * x[2];
}

// What slope_limiting_kernel decides for one element: may act (element and neighbours wet), the
// variables it writes back, and the 2-norm of each variable's vertex-value change.
struct SlOutcome {
    bool ok;
    unsigned mask;
    double del[3];
};

// WRITE: the limiter itself (writes the limited modes, the flags, the NaN scan); otherwise a dry
// run that only reports the decision (the learned indicator's judge and label pass).
template <typename Real, int P, bool WRITE>
RKDG_HD RKDG_FORCE_INLINE inline SlOutcome sl_limit_core(const LimiterArgs<Real>& L, int e) {
#pragma clang fp contract(off)
    constexpr int ND = Sizes<P>::ND;
    const int ne = L.ne;
    SlOutcome out{false, 0u, {0.0, 0.0, 0.0}};
    const bool wet = L.wet[e] != 0;
    bool wet_neigh = true;
    int nbr[3];
    for (int k = 0; k < 3; ++k) {
        nbr[k] = L.neighbor[(size_t)k * ne + e];
        const bool w = nbr[k] >= 0 ? (L.wet[nbr[k]] != 0) : wet;
        wet_neigh = wet_neigh && w;
    }
    uint8_t lim[3] = {0, 0, 0};
    out.ok = wet && wet_neigh;
    if (wet && wet_neigh) {
        using std::copysign;
        using std::fabs;
        using std::signbit;
        using std::sqrt;
        double ql[3][3];
        load_linear<Real, P>(L, e, ql);
        double bc[3], bcn[3][3];
        for (int k = 0; k < 3; ++k) bc[k] = L.baryctr[(size_t)k * ne + e];
        for (int f = 0; f < 3; ++f)
            for (int k = 0; k < 3; ++k) bcn[f][k] = nbr[f] >= 0 ? L.baryctr[(size_t)k * ne + nbr[f]] : bc[k];
        const double chi[3][3] = {{0.0, 0.5, 0.5}, {0.5, 0.0, 0.5}, {0.5, 0.5, 0.0}};
        const double h = bc[0] + L.bath_bc[e];
        const double u = bc[1] / h, v = bc[2] / h;
        const double c = sqrt(L.g * h);
        double delta[3][3];  // [var][bound]
        for (int b = 0; b < 3; ++b) {
            const int e1 = b, e2 = (b + 1) % 3;
            const double mx = L.median[(size_t)(2 * b) * ne + e], my = L.median[(size_t)(2 * b + 1) * ne + e];
            // R and invR along the median (problem_jacobian/swe_jacobian.hpp:462-480, 531-550)
            const double R[3][3] = {{1.0, 0.0, 1.0}, {u - c * mx, -my, u + c * mx}, {v - c * my, mx, v + c * my}};
            const double un = u * mx + v * my;
            const double iR[3][3] = {{(c + un) / (2 * c), -mx / (2 * c), -my / (2 * c)},
                                     {u * my - v * mx, -my, mx},
                                     {(c - un) / (2 * c), mx / (2 * c), my / (2 * c)}};
            double qm[3];
            for (int k = 0; k < 3; ++k) qm[k] = (ql[k][0] * chi[0][b] + ql[k][1] * chi[1][b]) + ql[k][2] * chi[2][b];
            double wm[3], w0[3], w1[3], w2[3];
            mv3(iR, qm, wm);
            mv3(iR, bc, w0);
            mv3(iR, bcn[e1], w1);
            mv3(iR, bcn[e2], w2);
            const double a0 = L.alpha[(size_t)(2 * b) * ne + e], a1 = L.alpha[(size_t)(2 * b + 1) * ne + e];
            const double Mr2 = L.M * L.rsq[(size_t)b * ne + e];
            double dc[3];
            for (int k = 0; k < 3; ++k) {
                const double wt = wm[k] - w0[k];
                const double wd = a0 * (w1[k] - w0[k]) + a1 * (w2[k] - w0[k]);
                if (fabs(wt) <= Mr2)
                    dc[k] = wt;
                else if (signbit(wt) == signbit(wd))
                    dc[k] = copysign(1.0, wt) * rmin(fabs(wt), fabs(L.nu * wd));
                else
                    dc[k] = 0.0;
            }
            double col[3];
            mv3(R, dc, col);
            for (int k = 0; k < 3; ++k) delta[k][b] = col[k];
        }
        for (int k = 0; k < 3; ++k) {
            double sum = 0.0;
            for (int b = 0; b < 3; ++b) sum += delta[k][b];
            if (sum != 0.0) {
                double pos = 0.0, neg = 0.0;
                for (int b = 0; b < 3; ++b) {
                    pos += rmax(0.0, delta[k][b]);
                    neg += rmax(0.0, -delta[k][b]);
                }
                const double tp = rmin(1.0, neg / pos);
                const double tn = rmin(1.0, pos / neg);
                for (int b = 0; b < 3; ++b) delta[k][b] = tp * rmax(0.0, delta[k][b]) - tn * rmax(0.0, -delta[k][b]);
            }
        }
        const double Tm[3][3] = {{-1.0, 1.0, 1.0}, {1.0, -1.0, 1.0}, {1.0, 1.0, -1.0}};
        for (int k = 0; k < 3; ++k) {
            double qv[3];
            for (int vv = 0; vv < 3; ++vv)
                qv[vv] = bc[k] + ((delta[k][0] * Tm[0][vv] + delta[k][1] * Tm[1][vv]) + delta[k][2] * Tm[2][vv]);
            const double d0 = qv[0] - ql[k][0], d1 = qv[1] - ql[k][1], d2 = qv[2] - ql[k][2];
            const double del_norm = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
            const double q_norm = sqrt((ql[k][0] * ql[k][0] + ql[k][1] * ql[k][1]) + ql[k][2] * ql[k][2]);
            out.del[k] = del_norm;
            if (del_norm / q_norm > 1.0e-12) {
                lim[k] = 1;
                out.mask |= 1u << k;
                if constexpr (WRITE) {
                    double m[3];
                    from_vertices(qv, m);
                    for (int d = 0; d < 3; ++d) L.q[(size_t)(k * ND + d) * ne + e] = (Real)m[d];
                    for (int d = 3; d < ND; ++d) L.q[(size_t)(k * ND + d) * ne + e] = Real(0.0);
                }
            }
        }
    }
    if constexpr (!WRITE) return out;
    for (int k = 0; k < 3; ++k) L.limited[(size_t)k * ne + e] = lim[k];
    if (L.nan_check) {
        bool has_nan = false;
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < ND; ++d) {
                const Real x = L.q[(size_t)(k * ND + d) * ne + e];
                has_nan |= (x != x);
            }
        if (has_nan) {
            const unsigned long long key = ((L.step_ctr ? *L.step_ctr * (unsigned long long)L.key_ns + L.stage_key : L.stage_key) << 32) | (unsigned long long)(unsigned)e;
#ifdef __HIP_DEVICE_COMPILE__
            atomicMin(L.nan_key, key);
#else
            if (key < *L.nan_key) *L.nan_key = key;
#endif
        }
    }
    return out;
}

template <typename Real, int P>
RKDG_HD RKDG_FORCE_INLINE inline void sl_limit_element(const LimiterArgs<Real>& L, const FaceTables<Real>&, int e) {
    (void)sl_limit_core<Real, P, true>(L, e);
}

// Relative size of a write-back for the indicator's labels: the change of each variable's vertex
// values over that variable's physical scale (h_s for ze, h_s sqrt(g h_s) for qx, qy, with
// h_s = max(ze mode 0 + bathymetry at the barycentre, 0.01)); the largest over written-back variables.
template <typename Real>
RKDG_HD inline double sl_change(const LimiterArgs<Real>& L, const SlOutcome& o, int e, double h_floor) {
#pragma clang fp contract(off)
    double h = (double)L.q[e] + L.bath_bc[e];
    h = h > h_floor ? h : h_floor;
    using std::sqrt;
    const double sq = h * sqrt(L.g * h);
    double r = 0.0;
    for (int k = 0; k < 3; ++k)
        if (o.mask & (1u << k)) {
            const double c = o.del[k] / (k == 0 ? h : sq);
            r = c > r ? c : r;
        }
    return r;
}

}  // namespace rkdg
