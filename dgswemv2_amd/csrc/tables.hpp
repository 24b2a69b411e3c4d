// Reference-element tables of the RKDG engine (host, FP64).
//
// Dubiner basis and Jacobi recurrences follow basis/bases_2D/basis_dubiner_2D.cpp:4-257 and
// basis/polynomials/basis_polynomials.cpp:4-67 operation by operation, so the tables the
// device kernels load are bit-identical to the reference's.  Quadrature data comes from
// quadrature_data.h (Dunavant rules, Gauss-Legendre rules).
#pragma once

#include <array>
#include <vector>

namespace rkdg {

struct Rule2 {
    std::vector<double> w;                     // weights
    std::vector<std::array<double, 2>> z;      // points on the (-1,-1),(1,-1),(-1,1) triangle
};
struct Rule1 {
    std::vector<double> w;
    std::vector<double> z;
};

// Dunavant rule of polynomial order p (integration_dunavant_2D.cpp:4-76), expanded from the
// orbit generators in the reference's point order.
Rule2 dunavant_rule(int p);
// Gauss-Legendre rule exact to order p: ceil((p+1)/2) points (integration_gausslegendre_1D.cpp:4-30).
Rule1 gauss_legendre_rule(int p);

// Dubiner modes at points: phi[dof*npt + pt], dz1/dz2 likewise (any of them may be null).
void dubiner_eval(int p, const std::vector<std::array<double, 2>>& pts, double* phi, double* dz1, double* dz2);
// Mode index -> (p, q) triangular index (basis_dubiner_2D.cpp:18-23).
void dubiner_pq(int dof, int& p, int& q);

// Master triangle for order p (master_triangle.tpp:5-87): rule 2p, phi/dphi at the rule points,
// int_phi_fact = phi^T w, int_dphi_fact = dphi^T w, diagonal M^-1, edge tables at the
// Gauss-Legendre rule 2p+1 for each local edge (BoundaryToMasterCoordinates :89-122).
struct Master {
    int p = 0, ndof = 0, ngp = 0, ngp_e = 0;
    Rule2 rule;
    Rule1 erule;
    std::vector<double> phi;        // [ndof][ngp]
    std::vector<double> dphi[2];    // [ndof][ngp]
    std::vector<double> ipf;        // [ngp][ndof]  phi(d,g)*w(g)
    std::vector<double> idphi[2];   // [ngp][ndof]  dphi(d,g)*w(g)
    std::vector<double> minv;       // [ndof]
    std::vector<double> phi_e[3];   // [ndof][ngp_e] per local edge
    std::vector<std::array<double, 2>> zedge[3];  // master coordinates of the edge points
};
Master make_master(int p);

}  // namespace rkdg
