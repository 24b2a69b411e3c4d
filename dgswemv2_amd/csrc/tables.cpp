// Reference-element tables (host, FP64).  See tables.hpp.
#include "tables.hpp"

#include <cmath>
#include <stdexcept>

#include "quadrature_data.h"

namespace rkdg {

Rule2 dunavant_rule(int p) {
    if (p < 0 || p > 20) throw std::invalid_argument("Dunavant order outside 0..20");
    const int order = p == 0 ? 1 : p;  // rules 0 and 1 coincide (integration_dunavant_2D.cpp:128-131)
    const int first = rkdg_quad::kDunavantIndex[order - 1][0];
    const int count = rkdg_quad::kDunavantIndex[order - 1][1];
    // Orbit expansion order of GetRule (:30-72): identity, then the cyclic shifts, then
    // (for 6-point orbits) the three transposed pairs.
    static const int kPairs[6][2] = {{0, 1}, {1, 2}, {2, 0}, {1, 0}, {2, 1}, {0, 2}};
    Rule2 r;
    for (int o = first; o < first + count; ++o) {
        const rkdg_quad::DunavantOrbit& g = rkdg_quad::kDunavantOrbits[o];
        const double lam[3] = {g.l0, g.l1, g.l2};
        for (int k = 0; k < g.orbit; ++k) {
            r.w.push_back(2 * g.w);
            r.z.push_back({2 * lam[kPairs[k][0]] - 1, 2 * lam[kPairs[k][1]] - 1});
        }
    }
    return r;
}

Rule1 gauss_legendre_rule(int p) {
    const int n = (int)std::ceil(((double)p + 1) / 2);
    if (n < 1 || n > 33) throw std::invalid_argument("Gauss-Legendre point count outside 1..33");
    const int first = rkdg_quad::kGaussLegendreIndex[n - 1];
    Rule1 r;
    for (int i = 0; i < n; ++i) {
        r.w.push_back(rkdg_quad::kGaussLegendre[first + i].w);
        r.z.push_back(rkdg_quad::kGaussLegendre[first + i].z);
    }
    return r;
}

// Jacobi P_n^(a,b) at the points: three-term recurrence with a 3-slot rotating buffer.
static void jacobi_values(unsigned n, unsigned a, unsigned b, const std::vector<double>& x, std::vector<double>& out) {
    const size_t m = x.size();
    out.assign(m, 0.0);
    for (size_t k = 0; k < m; ++k) {
        const double p1 = ((double)(a - b) + (double)(a + b + 2) * x[k]) / 2;
        if (n == 0) {
            out[k] = 1;
            continue;
        }
        if (n == 1) {
            out[k] = p1;
            continue;
        }
        double slot[3] = {1.0, p1, 0.0};
        for (unsigned i = 1; i < n; ++i) {
            const double c1 = (double)(2 * (i + 1) * (i + a + b + 1) * (2 * i + a + b));
            const double c2 = (double)((2 * i + a + b + 1) * (a * a - b * b));
            const double c3 = (double)((2 * i + a + b) * (2 * i + a + b + 1) * (2 * i + a + b + 2));
            const double c4 = (double)(2 * (i + a) * (i + b) * (2 * i + a + b + 2));
            slot[(i + 1) % 3] = ((c2 + c3 * x[k]) * slot[i % 3] - c4 * slot[(i - 1) % 3]) / c1;
        }
        out[k] = slot[n % 3];
    }
}

static void jacobi_deriv(unsigned n, unsigned a, unsigned b, const std::vector<double>& x, std::vector<double>& out) {
    if (n == 0) {
        out.assign(x.size(), 0.0);
    } else if (n == 1) {
        out.assign(x.size(), (a + b + n + 1) / 2.0);
    } else {
        jacobi_values(n - 1, a + 1, b + 1, x, out);
        for (double& v : out) v *= (a + b + n + 1) / 2.0;
    }
}

void dubiner_pq(int dof, int& p, int& q) {
    const unsigned t = (unsigned)std::ceil((-3. + std::sqrt(1. + 8. * (dof + 1))) / 2.);
    const unsigned lower = (t + 1) * t / 2;
    p = (int)(dof - lower);
    q = (int)(t - p);
}

// Derivative values at the collapsed vertex (-1, 1) (basis_dubiner_2D.cpp:177-211), as the
// second component of a two-term recursion in q.
static double singular_dz1(unsigned q) {
    double s0 = 3., s1 = 1.;
    for (unsigned k = 1; k <= q; ++k) {
        const double n0 = s0 + k + 2, n1 = s0 + s1;
        s0 = n0, s1 = n1;
    }
    return s1;
}
static double singular_dz2(unsigned q) {
    double s0 = 0.5, s1 = 0;
    for (unsigned k = 1; k <= q; ++k) {
        const double n0 = s0 + 0.5 * (k + 1), n1 = s0 + s1;
        s0 = n0, s1 = n1;
    }
    return s1;
}

void dubiner_eval(int P, const std::vector<std::array<double, 2>>& pts, double* phi, double* dz1, double* dz2) {
    const int ndof = (P + 1) * (P + 2) / 2;
    const size_t npt = pts.size();
    std::vector<double> n1(npt), n2(npt);
    for (size_t k = 0; k < npt; ++k) {
        n1[k] = 2 * (1 + pts[k][0]) / (1 - pts[k][1]) - 1;
        n2[k] = pts[k][1];
    }
    std::vector<double> psi_p, psi_pq, dpsi_p, dpsi_pq;
    for (int dof = 0; dof < ndof; ++dof) {
        int pi, qi;
        dubiner_pq(dof, pi, qi);
        const unsigned p = (unsigned)pi, q = (unsigned)qi;
        jacobi_values(p, 0, 0, n1, psi_p);
        jacobi_values(q, 2 * p + 1, 0, n2, psi_pq);
        jacobi_deriv(p, 0, 0, n1, dpsi_p);
        jacobi_deriv(q, 2 * p + 1, 0, n2, dpsi_pq);
        for (size_t k = 0; k < npt; ++k) {
            const bool collapsed = std::isnan(n1[k]);
            const double s = (1.0 - n2[k]) / 2.0;
            const double s_p = std::pow(s, (double)(int)p);
            if (phi) {
                double v = psi_p[k] * std::pow((1 - n2[k]) / 2, (double)(int)p) * psi_pq[k];
                if (collapsed) v = p == 0 ? (double)(q + 1) : 0.0;
                phi[dof * npt + k] = v;
            }
            if (dz1) {
                double v = (2.0 / (1.0 - n2[k])) * dpsi_p[k] * s_p * psi_pq[k];
                if (collapsed) v = p == 1 ? singular_dz1(q) : 0.0;
                dz1[dof * npt + k] = v;
            }
            if (dz2) {
                double v = ((1 + n1[k]) / (1.0 - n2[k])) * dpsi_p[k] * s_p * psi_pq[k] +
                           psi_p[k] * (s_p * dpsi_pq[k] - (p / 2.0) * std::pow(s, (double)(int)(p - 1)) * psi_pq[k]);
                if (collapsed) v = p == 0 ? 3 * singular_dz2(q) : (p == 1 ? singular_dz2(q + 1) : 0.0);
                dz2[dof * npt + k] = v;
            }
        }
    }
}

Master make_master(int p) {
    Master m;
    m.p = p;
    m.ndof = (p + 1) * (p + 2) / 2;
    m.rule = dunavant_rule(2 * p);
    m.ngp = (int)m.rule.w.size();
    m.erule = gauss_legendre_rule(2 * p + 1);
    m.ngp_e = (int)m.erule.w.size();
    const int nd = m.ndof, ng = m.ngp;
    m.phi.resize(nd * ng);
    m.dphi[0].resize(nd * ng);
    m.dphi[1].resize(nd * ng);
    dubiner_eval(p, m.rule.z, m.phi.data(), m.dphi[0].data(), m.dphi[1].data());
    m.ipf.resize(ng * nd);
    m.idphi[0].resize(ng * nd);
    m.idphi[1].resize(ng * nd);
    for (int d = 0; d < nd; ++d)
        for (int g = 0; g < ng; ++g) {
            m.ipf[g * nd + d] = m.phi[d * ng + g] * m.rule.w[g];
            m.idphi[0][g * nd + d] = m.dphi[0][d * ng + g] * m.rule.w[g];
            m.idphi[1][g * nd + d] = m.dphi[1][d * ng + g] * m.rule.w[g];
        }
    m.minv.resize(nd);
    for (int d = 0; d < nd; ++d) {
        int pp, qq;
        dubiner_pq(d, pp, qq);
        m.minv[d] = ((2 * pp + 1) * (pp + qq + 1) / 2.0);
    }
    for (int f = 0; f < 3; ++f) {
        std::vector<std::array<double, 2>>& zm = m.zedge[f];
        zm.resize(m.ngp_e);
        for (int g = 0; g < m.ngp_e; ++g) {
            const double z = m.erule.z[g];
            if (f == 0)
                zm[g] = {-z, z};
            else if (f == 1)
                zm[g] = {-1.0, -z};
            else
                zm[g] = {z, -1.0};
        }
        m.phi_e[f].resize(nd * m.ngp_e);
        dubiner_eval(p, zm, m.phi_e[f].data(), nullptr, nullptr);
    }
    return m;
}

}  // namespace rkdg
