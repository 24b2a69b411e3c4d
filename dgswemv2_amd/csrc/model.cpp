// Host model construction: see model.hpp.
#include "model.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <unordered_map>

#include "engine.hpp"
#include "ssprk_data.h"

namespace rkdg {

// ---------------------------------------------------------------------------------------
bool Stepper::init(int ns, int ord, double dt_, double ramp_days) {
    if (ns < 1 || ns > 8 || ord < 1) return false;  // the published schemes have 1-8 stages
    nstages = ns, order = ord, dt = dt_, ramp_duration = ramp_days;
    step = 0, stage = 0, t = 0;
    const bool no_ramp = std::abs(ramp_duration) < std::numeric_limits<double>::epsilon() * 100;
    ramp = no_ramp ? 1. : 0.;
    ark.assign(ns, std::vector<double>(ns, 0.0));
    brk.assign(ns, std::vector<double>(ns, 0.0));
    bool found = false;
    for (const auto& en : rkdg_ssprk::kEntries)
        if (en.stages == ns && en.order == ord) {
            ark[en.row][en.col] = en.alpha;
            brk[en.row][en.col] = en.beta;
            found = true;
        }
    if (!found) return false;
    // stage-time offsets (explicit_ssp_rk_stepper.cpp:259-276); crk's diagonal is never set,
    // so drk[k] sums crk[k-1][0..k-2] only
    std::vector<std::vector<double>> crk(ns, std::vector<double>(ns, 0.0));
    for (int i = 0; i < ns; ++i)
        for (int k = 0; k < i; ++k) {
            double casum = 0.;
            for (int l = k + 1; l < i; ++l) casum += crk[l][k] * ark[i][l];
            crk[i][k] = brk[i][k] + casum;
        }
    drk.assign(ns, 0.0);
    for (int k = 1; k < ns; ++k)
        for (int l = 0; l < k; ++l) drk[k] = drk[k] + crk[k - 1][l];
    return true;
}

void Stepper::advance() {
    ++stage;
    stage = stage % nstages;
    if (stage == 0) {
        t += dt;
        ++step;
    }
    if (!(std::abs(ramp_duration) < std::numeric_limits<double>::epsilon() * 100))
        ramp = std::tanh(2 * (stage_time() / 86400) / ramp_duration);
}

// ---------------------------------------------------------------------------------------
namespace {

constexpr int kNoNeighbour = -1;

void problem_ic_host(const rkdg_params& prm, double t, double x, double y, double* q) {
    problem_ic<double>(prm.problem, prm.problem_param, prm.g, t, x, y, q);
}

// C = A (ra x ca) * B (ca x cb), row-major, sums from the first product.
std::vector<double> matmul(const std::vector<double>& A, int ra, int ca, const std::vector<double>& B, int cb) {
    std::vector<double> C((size_t)ra * cb);
    for (int i = 0; i < ra; ++i)
        for (int j = 0; j < cb; ++j) {
            double s = A[(size_t)i * ca] * B[j];
            for (int k = 1; k < ca; ++k) s = s + A[(size_t)i * ca + k] * B[(size_t)k * cb + j];
            C[(size_t)i * cb + j] = s;
        }
    return C;
}

struct Tri {
    double x[3], y[3], z[3];
    void J(double M[2][2]) const {  // shape_straighttriangle.cpp:95-108
        M[0][0] = (x[1] - x[0]) / 2.0;
        M[0][1] = (x[2] - x[0]) / 2.0;
        M[1][0] = (y[1] - y[0]) / 2.0;
        M[1][1] = (y[2] - y[0]) / 2.0;
    }
    double det() const {
        double M[2][2];
        J(M);
        return M[0][0] * M[1][1] - M[1][0] * M[0][1];
    }
    double surfaceJ(int b) const {
        const double dx = x[(b + 2) % 3] - x[(b + 1) % 3], dy = y[(b + 2) % 3] - y[(b + 1) % 3];
        return std::sqrt(std::pow(dx, 2.0) + std::pow(dy, 2.0)) / 2.0;
    }
    void normal(int b, double& nx, double& ny) const {
        const double d = det();
        const double cw = d / std::abs(d);
        const double dx = x[(b + 2) % 3] - x[(b + 1) % 3], dy = y[(b + 2) % 3] - y[(b + 1) % 3];
        const double length = std::hypot(dx, dy);
        nx = cw * dy / length;
        ny = -cw * dx / length;
    }
};

// shape functions at local points: P[node][pt] (GetPsi, shape_straighttriangle.cpp:170-182)
std::vector<double> shape_psi(const std::vector<std::array<double, 2>>& pts) {
    const size_t n = pts.size();
    std::vector<double> P(3 * n);
    for (size_t k = 0; k < n; ++k) {
        P[k] = -(pts[k][0] + pts[k][1]) / 2;
        P[n + k] = (1 + pts[k][0]) / 2;
        P[2 * n + k] = (1 + pts[k][1]) / 2;
    }
    return P;
}

}  // namespace

std::string build_model(const rkdg_mesh_desc& md, const rkdg_params& prm, Model& M, bool reference_face_order) {
    M.reference_face_order = reference_face_order;
    if (md.n_elements <= 0 || md.n_nodes <= 0 || !md.node_xyz || !md.elt_nodes) return "empty mesh";
    if (md.n_bound_edges < 0 || (md.n_bound_edges > 0 && (!md.bound_edge_nodes || !md.bound_edge_type)))
        return "boundary edge count negative or its arrays missing";
    if (md.n_constituents < 0) return "negative constituent count";
    if (md.n_constituents > 0 && (!md.frequency || !md.forcing_fact || !md.equilib_arg))
        return "constituent arrays missing";
    if (md.n_levee_pairs < 0 || (md.n_levee_pairs > 0 && (!md.levee_nodes || !md.levee_param)))
        return "levee pair count negative or its arrays missing";
    if (prm.problem < RKDG_PROBLEM_NONE || prm.problem > RKDG_PROBLEM_DAM_BREAK) return "unknown problem";
    if (prm.friction < RKDG_FRICTION_NONE || prm.friction > RKDG_FRICTION_MANNING) return "unknown friction law";
    if (prm.p < 1 || prm.p > 3) return "polynomial order outside the device-built set {1, 2, 3}";
    // the kernels address every element-fastest array with 32-bit byte offsets
    if ((size_t)md.n_elements * 3 * 6 * sizeof(double) >= ((size_t)1 << 32))
        return "mesh too large: state arrays exceed 4 GiB of 32-bit buffer offsets";
    if (!M.stepper.init(prm.nstages, prm.order, prm.dt, prm.ramp_duration)) return "invalid Runge-Kutta method";
    M.prm = prm;
    M.p = prm.p;
    M.master = make_master(prm.p);
    M.ndof = M.master.ndof, M.ngp = M.master.ngp, M.ngpe = M.master.ngp_e;
    const int ne = md.n_elements, nd = M.ndof, ng = M.ngp, nge = M.ngpe;
    M.ne = ne;
    if (prm.wetting_drying) {  // swe_pre_init_prob_params.hpp:33-37
        M.h_o = prm.h_o;
        M.h_thr = 1.0e6 * M.h_o * std::numeric_limits<double>::epsilon();
    }
    if (prm.slope_limiting) M.sl_M = prm.sl_M, M.sl_nu = prm.sl_nu;
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) {
            const int v = md.elt_nodes[3 * e + k];
            if (v < 0 || v >= md.n_nodes) return "element node index out of range";
        }
    for (size_t i = 0; i < 3 * (size_t)md.n_nodes; ++i)
        if (!std::isfinite(md.node_xyz[i])) return "non-finite node coordinate or depth";
    for (int e = 0; e < ne; ++e) {  // collinear vertices (a repeated node included): no Jacobian inverse
        const double* a = md.node_xyz + 3 * (size_t)md.elt_nodes[3 * e];
        const double* b = md.node_xyz + 3 * (size_t)md.elt_nodes[3 * e + 1];
        const double* c = md.node_xyz + 3 * (size_t)md.elt_nodes[3 * e + 2];
        if ((b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]) == 0.0) return "degenerate element (zero area)";
    }

    // ---- edges by node pair (mesh_metadata.cpp:31-85)
    std::map<std::pair<int, int>, uint8_t> btype_of;
    for (int i = 0; i < md.n_bound_edges; ++i) {
        const int a = md.bound_edge_nodes[2 * i], b = md.bound_edge_nodes[2 * i + 1];
        btype_of[{std::min(a, b), std::max(a, b)}] = md.bound_edge_type[i];
    }
    std::vector<int> nbr(3 * (size_t)ne, kNoNeighbour), nbr_face(3 * (size_t)ne, 0);
    std::vector<int> btype(3 * (size_t)ne, -1);  // BC code of boundary faces
    std::vector<uint8_t> levee(3 * (size_t)ne, 0);  // face is a levee (barrier) interface
    std::map<std::pair<int, int>, int> levee_pair;  // (min, max) node pair -> levee_nodes row
    {
        std::unordered_map<uint64_t, std::array<int, 4>> dict;
        dict.reserve((size_t)ne * 2);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                const uint32_t a = md.elt_nodes[3 * e + (k + 1) % 3], b = md.elt_nodes[3 * e + (k + 2) % 3];
                const uint64_t key = (uint64_t)std::min(a, b) << 32 | std::max(a, b);
                auto it = dict.find(key);
                if (it == dict.end())
                    dict.emplace(key, std::array<int, 4>{e, k, -1, 0});
                else if (it->second[2] < 0)
                    it->second[2] = e, it->second[3] = k;
                else
                    return "edge shared by more than two elements";
            }
        for (auto& kv : dict) {
            const auto& v = kv.second;
            if (v[2] >= 0) {
                nbr[3 * v[0] + v[1]] = v[2], nbr_face[3 * v[0] + v[1]] = v[3];
                nbr[3 * v[2] + v[3]] = v[0], nbr_face[3 * v[2] + v[3]] = v[1];
            } else {
                const int a = (int)(kv.first >> 32), b = (int)(kv.first & 0xffffffffu);
                auto it = btype_of.find({a, b});
                if (it == btype_of.end()) return "boundary edge without a boundary type";
                if (it->second > RKDG_BC_FUNCTION && it->second != RKDG_BC_LEVEE) return "unsupported boundary type";
                btype[3 * v[0] + v[1]] = it->second;
            }
        }
        // levee faces: the element across the barrier owns the partner edge, whose nodes are the
        // barrier partners of this edge's nodes (AdcircFormat::get_internal_node_pair,
        // adcirc_format.cpp:260-310; MeshMetaData, mesh_metadata.cpp:74-80)
        std::unordered_map<int, int> partner;
        for (int i = 0; i < md.n_levee_pairs; ++i) {
            const int f = md.levee_nodes[2 * i], b = md.levee_nodes[2 * i + 1];
            if (f < 0 || f >= md.n_nodes || b < 0 || b >= md.n_nodes) return "levee node out of range";
            if (partner.count(f) || partner.count(b)) return "levee node in more than one pair";
            partner[f] = b, partner[b] = f;
            levee_pair[{std::min(f, b), std::max(f, b)}] = i;
        }
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                if (btype[3 * e + k] != RKDG_BC_LEVEE) continue;
                const int a = md.elt_nodes[3 * e + (k + 1) % 3], b = md.elt_nodes[3 * e + (k + 2) % 3];
                auto pa = partner.find(a), pb = partner.find(b);
                if (pa == partner.end() || pb == partner.end()) return "levee edge without barrier node pairs";
                const uint32_t x = pa->second, y = pb->second;
                auto it = dict.find((uint64_t)std::min(x, y) << 32 | std::max(x, y));
                if (it == dict.end() || it->second[2] >= 0) return "levee edge without a single partner element";
                nbr[3 * e + k] = it->second[0], nbr_face[3 * e + k] = it->second[1];
                levee[3 * e + k] = 1;
                btype[3 * e + k] = -1;
            }
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k)
                if (levee[3 * e + k] && (nbr[3 * nbr[3 * e + k] + nbr_face[3 * e + k]] != e))
                    return "levee partner edges do not pair up";
    }

    // ---- spherical meshes: (longitude, latitude) in degrees projected onto a cylinder before
    // anything else (preprocess_mesh_data, swe_pre_preprocess_mesh.hpp:5-27)
    std::vector<double> proj;
    const double* xyz = md.node_xyz;
    if (prm.spherical_mesh) {
        if (!(prm.sph_R > 0)) return "spherical mesh needs a positive radius";
        const double R = prm.sph_R, R_o = R * std::cos(prm.sph_lat_o * kPI / 180.0);
        proj.assign(md.node_xyz, md.node_xyz + 3 * (size_t)md.n_nodes);
        for (int v = 0; v < md.n_nodes; ++v) {
            proj[3 * (size_t)v] = R_o * (proj[3 * (size_t)v] - prm.sph_lon_o) * kPI / 180.0;
            proj[3 * (size_t)v + 1] = R * proj[3 * (size_t)v + 1] * kPI / 180.0;
        }
        xyz = proj.data();
    }
    if (prm.sp_correction && !(prm.sph_R > 0)) return "spherical correction needs a positive radius";

    // ---- element geometry (Element::Initialize, element.hpp:166-272)
    std::vector<Tri> tri(ne);
    M.geom.assign(5 * (size_t)ne, 0.0);
    M.vxy.assign(6 * (size_t)ne, 0.0);
    for (int e = 0; e < ne; ++e) {
        Tri& T = tri[e];
        for (int k = 0; k < 3; ++k) {
            const int v = md.elt_nodes[3 * e + k];
            T.x[k] = xyz[3 * v], T.y[k] = xyz[3 * v + 1], T.z[k] = xyz[3 * v + 2];
        }
        if (T.det() < 0) {  // coordinates-only CCW swap (shape_straighttriangle.cpp:9-11)
            std::swap(T.x[0], T.x[2]);
            std::swap(T.y[0], T.y[2]);
            std::swap(T.z[0], T.z[2]);
        }
        double Jm[2][2];
        T.J(Jm);
        const double det = Jm[0][0] * Jm[1][1] - Jm[1][0] * Jm[0][1];
        const double invdet = 1.0 / det;  // Eigen 2x2 inverse
        M.geom[e] = Jm[1][1] * invdet;                    // Jinv(0,0)
        M.geom[(size_t)ne + e] = -Jm[0][1] * invdet;      // Jinv(0,1)
        M.geom[2 * (size_t)ne + e] = -Jm[1][0] * invdet;  // Jinv(1,0)
        M.geom[3 * (size_t)ne + e] = Jm[0][0] * invdet;   // Jinv(1,1)
        M.geom[4 * (size_t)ne + e] = std::abs(det);
        for (int k = 0; k < 3; ++k) M.vxy[(size_t)k * ne + e] = T.x[k], M.vxy[(size_t)(3 + k) * ne + e] = T.y[k];
    }

    // ---- edge records: interfaces in (in, ex) order, then boundaries by (type, element, face).
    // Interfaces, levees included, carry the in-side normal and surface Jacobian, which the ex
    // side uses negated (interface.hpp:206-217).  A levee edge also has a record of barrier
    // height and weir coefficients at the in-side Gauss points (ISP::Levee::Initialize,
    // rkdg_swe_is_levee.hpp:50-60: nodal values times GetBoundaryPsi).
    std::vector<int> face_edge(3 * (size_t)ne, -1);
    std::vector<double> eg[3];
    M.eforce.clear();
    for (int e = 0; e < ne; ++e) {
        std::vector<std::pair<int, int>> up;
        for (int k = 0; k < 3; ++k)
            if (nbr[3 * e + k] > e) up.push_back({nbr[3 * e + k], k});
        std::sort(up.begin(), up.end());
        for (auto& pr : up) {
            const int id = (int)eg[0].size(), k = pr.second;
            double nx, ny;
            tri[e].normal(k, nx, ny);
            eg[0].push_back(nx), eg[1].push_back(ny), eg[2].push_back(tri[e].surfaceJ(k));
            face_edge[3 * e + k] = id;
            face_edge[3 * pr.first + nbr_face[3 * e + k]] = id;
            int rec = -1;
            if (levee[3 * e + k]) {
                // in-side node i pairs with the ex-side face node 1 - i (swe_pre_create_intface.hpp:50-69)
                const int x = pr.first, kx = nbr_face[3 * e + k];
                const int in_n[2] = {md.elt_nodes[3 * e + (k + 1) % 3], md.elt_nodes[3 * e + (k + 2) % 3]};
                const int ex_n[2] = {md.elt_nodes[3 * x + (kx + 1) % 3], md.elt_nodes[3 * x + (kx + 2) % 3]};
                double par[2][3];
                for (int i = 0; i < 2; ++i) {
                    const int a = in_n[i], b = ex_n[1 - i];
                    auto it = levee_pair.find({std::min(a, b), std::max(a, b)});
                    if (it == levee_pair.end()) return "unable to find levee data for a barrier edge";
                    for (int c = 0; c < 3; ++c) par[i][c] = md.levee_param[3 * (size_t)it->second + c];
                }
                rec = (int)(M.levee_gp.size() / (3 * (size_t)nge));
                for (int c = 0; c < 3; ++c)
                    for (int g = 0; g < nge; ++g) {
                        const double z = M.master.erule.z[g];
                        M.levee_gp.push_back(par[0][c] * ((1 - z) / 2) + par[1][c] * ((1 + z) / 2));
                    }
            }
            M.eforce.push_back(rec);
        }
    }
    M.nintf = (int)eg[0].size();
    std::vector<std::array<int, 3>> bedges;  // (type, element, face) in reference order
    for (int ty = 0; ty < 4; ++ty)
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k)
                if (btype[3 * e + k] == ty) bedges.push_back({ty, e, k});
    M.ncon = md.n_constituents;
    for (int c = 0; c < M.ncon; ++c) {
        M.freq.push_back(md.frequency[c]);
        M.ffact.push_back(md.forcing_fact[c]);
        M.eqarg.push_back(md.equilib_arg[c]);
    }
    for (auto& b : bedges) {
        const int id = (int)eg[0].size(), e = b[1], k = b[2];
        double nx, ny;
        tri[e].normal(k, nx, ny);
        eg[0].push_back(nx), eg[1].push_back(ny), eg[2].push_back(tri[e].surfaceJ(k));
        face_edge[3 * e + k] = id;
        int fr = -1;
        if (b[0] == RKDG_BC_TIDE || b[0] == RKDG_BC_FLOW) {
            if (M.ncon == 0 || !md.node_amplitude || !md.node_phase) return "tide/flow boundary without forcing data";
            fr = (int)(M.famp.size() / ((size_t)M.ncon * nge));
            // nodal amplitude/phase at the edge's two nodes interpolated to the Gauss points
            // (Tide::Initialize, rkdg_swe_bc_tide.hpp:54-68; GetBoundaryPsi)
            const int n0 = md.elt_nodes[3 * e + (k + 1) % 3], n1 = md.elt_nodes[3 * e + (k + 2) % 3];
            for (int c = 0; c < M.ncon; ++c)
                for (int g = 0; g < nge; ++g) {
                    const double z = M.master.erule.z[g];
                    const double N1 = (1 - z) / 2, N2 = (1 + z) / 2;
                    M.famp.push_back(md.node_amplitude[(size_t)n0 * M.ncon + c] * N1 +
                                     md.node_amplitude[(size_t)n1 * M.ncon + c] * N2);
                    M.fphase.push_back(md.node_phase[(size_t)n0 * M.ncon + c] * N1 +
                                       md.node_phase[(size_t)n1 * M.ncon + c] * N2);
                }
        }
        M.eforce.push_back(fr);
    }
    M.nedge = (int)eg[0].size();
    M.closed_forms = prm.function_source != 0;
    for (auto& b : bedges) M.closed_forms |= (b[0] == RKDG_BC_FUNCTION);

    // ---- spherical correction sp = cos(lat_o) / cos(y / R) at the volume points and at each
    // edge's in-side points (swe_pre_init_data.hpp:31-48, 80-160), y interpolated from the
    // element's (CCW-fixed) vertices with the linear shape functions, (y0 psi0 + y1 psi1) + y2 psi2.
    // Built into the kernels with the closed-form problem functions.
    if (prm.sp_correction) {
        const double cphi = std::cos(prm.sph_lat_o * kPI / 180.0), R = prm.sph_R;
        auto y_at = [&](int e, double z1, double z2) {
            const Tri& T = tri[e];
            return (T.y[0] * (-(z1 + z2) / 2) + T.y[1] * ((1 + z1) / 2)) + T.y[2] * ((1 + z2) / 2);
        };
        M.sp_vol.assign((size_t)ng * ne, 1.0);
        for (int e = 0; e < ne; ++e)
            for (int g = 0; g < ng; ++g)
                M.sp_vol[(size_t)g * ne + e] = cphi / std::cos(y_at(e, M.master.rule.z[g][0], M.master.rule.z[g][1]) / R);
        M.sp_edge.assign((size_t)nge * M.nedge, 1.0);
        for (int e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                const int id = face_edge[3 * e + k];
                const bool in_side = nbr[3 * e + k] < 0 || nbr[3 * e + k] > e;  // boundaries: the element
                if (id < 0 || !in_side) continue;
                for (int g = 0; g < nge; ++g)
                    M.sp_edge[(size_t)g * M.nedge + id] =
                        cphi / std::cos(y_at(e, M.master.zedge[k][g][0], M.master.zedge[k][g][1]) / R);
            }
        M.closed_forms = true;
    }
    M.egeom.resize(3 * (size_t)M.nedge);
    for (int c = 0; c < 3; ++c) std::copy(eg[c].begin(), eg[c].end(), M.egeom.begin() + (size_t)c * M.nedge);

    // ---- per-element face slots in rhs accumulation order
    M.nb.assign(3 * (size_t)ne, 0);
    M.edge.assign(3 * (size_t)ne, 0);
    M.meta.assign(ne, 0);
    M.neighbor.assign(3 * (size_t)ne, -1);
    M.sl_face_slot.assign(3 * (size_t)ne, 0);
    for (int e = 0; e < ne; ++e) {
        // (class, order key, face): internal interfaces by neighbour ID, then levee interfaces by
        // neighbour ID (the interface vectors run Internal then Levee), then boundaries
        std::array<std::array<int, 3>, 3> key;
        for (int k = 0; k < 3; ++k) {
            if (nbr[3 * e + k] >= 0)
                key[k] = {levee[3 * e + k] ? 1 : 0, nbr[3 * e + k], k};
            else
                key[k] = {2, btype[3 * e + k], k};
            M.neighbor[(size_t)k * ne + e] = nbr[3 * e + k];
        }
        if (reference_face_order) std::sort(key.begin(), key.end());
        uint32_t meta = 0;
        for (int s = 0; s < 3; ++s) {
            const int k = key[s][2];
            M.nb[(size_t)s * ne + e] = nbr[3 * e + k] >= 0 ? nbr[3 * e + k] : -1 - btype[3 * e + k];
            M.edge[(size_t)s * ne + e] = face_edge[3 * e + k];
            meta |= (uint32_t)k << (4 * s);
            meta |= (uint32_t)(nbr[3 * e + k] >= 0 ? nbr_face[3 * e + k] : 0) << (4 * s + 2);
            meta |= (uint32_t)levee[3 * e + k] << (12 + s);
            M.sl_face_slot[(size_t)k * ne + e] = s;
        }
        M.meta[e] = meta;
    }

    // ---- initialize_data_serial (swe_pre_init_data.hpp:8-327)
    const Master& m = M.master;
    const std::vector<double> psi = shape_psi(m.rule.z);
    M.bath.assign((size_t)nd * ne, 0.0);
    M.q0.assign(3 * (size_t)nd * ne, 0.0);
    M.wet0.assign(ne, 1);
    M.bath_vrtx.assign(3 * (size_t)ne, 0.0);
    M.gmn2.assign(ne, 0.0);
    // per element: v (r x 3 nodal or r x ng point values) -> * psi -> * int_phi_fact -> * M^-1,
    // each product summed from its first term (the order of the reference's Eigen products);
    // fixed-size arrays (p <= 3: ndof <= 10, ngp <= 12)
    constexpr int kMax = 12;
    if (nd > kMax || ng > kMax) return "element tables larger than the device-built set";
    double ipf[kMax * kMax], minv[kMax];
    auto project = [&](const double* v, int rows, bool nodal, double* out) {  // out[rows][nd]
        double at_gp[3][kMax], t[3][kMax];
        for (int r = 0; r < rows; ++r)
            for (int g = 0; g < ng; ++g) {
                if (!nodal) {
                    at_gp[r][g] = v[r * ng + g];
                    continue;
                }
                double s = v[r * 3 + 0] * psi[g];
                for (int k = 1; k < 3; ++k) s = s + v[r * 3 + k] * psi[(size_t)k * ng + g];
                at_gp[r][g] = s;
            }
        for (int r = 0; r < rows; ++r)
            for (int d = 0; d < nd; ++d) {
                double s = at_gp[r][0] * ipf[d];
                for (int g = 1; g < ng; ++g) s = s + at_gp[r][g] * ipf[g * nd + d];
                t[r][d] = s;
            }
        // times the diagonal M^-1 through a full product: the zero terms add exact zeros
        for (int r = 0; r < rows; ++r)
            for (int d = 0; d < nd; ++d) {
                double s = t[r][0] * (d == 0 ? minv[0] : 0.0);
                for (int k = 1; k < nd; ++k) s = s + t[r][k] * (k == d ? minv[k] : 0.0);
                out[r * nd + d] = s;
            }
    };
    for (int e = 0; e < ne; ++e) {
        const double adet = M.geom[4 * (size_t)ne + e];
        for (int g = 0; g < ng; ++g)
            for (int d = 0; d < nd; ++d) ipf[g * nd + d] = m.ipf[(size_t)g * nd + d] * adet;
        for (int d = 0; d < nd; ++d) minv[d] = m.minv[d] / adet;
        // bathymetry: L2ProjectionNode of nodal depths (element.hpp:317-321)
        double b[kMax], q[3 * kMax];
        project(tri[e].z, 1, true, b);
        for (int d = 0; d < nd; ++d) M.bath[(size_t)d * ne + e] = b[d];
        if (prm.ic_constant) {
            const double un[9] = {prm.ze0, prm.ze0, prm.ze0, prm.qx0, prm.qx0, prm.qx0, prm.qy0, prm.qy0, prm.qy0};
            project(un, 3, true, q);
        } else {  // L2ProjectionF of ic_q(0, x) (element.hpp:307-313)
            double f[3 * kMax];
            for (int g = 0; g < ng; ++g) {
                const double x = tri[e].x[0] * psi[g] + tri[e].x[1] * psi[ng + g] + tri[e].x[2] * psi[2 * ng + g];
                const double y = tri[e].y[0] * psi[g] + tri[e].y[1] * psi[ng + g] + tri[e].y[2] * psi[2 * ng + g];
                double v[3];
                problem_ic_host(prm, 0.0, x, y, v);
                f[g] = v[0], f[ng + g] = v[1], f[2 * ng + g] = v[2];
            }
            project(f, 3, false, q);
        }
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < nd; ++d) M.q0[(size_t)(k * nd + d) * ne + e] = q[k * nd + d];
        for (int v = 0; v < 3; ++v) M.bath_vrtx[(size_t)v * ne + e] = tri[e].z[v];
    }
    if (prm.friction == RKDG_FRICTION_MANNING) {  // :166-197
        if (!md.node_manning_n) return "Manning friction without nodal Manning n";
        for (int e = 0; e < ne; ++e) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += md.node_manning_n[md.elt_nodes[3 * e + k]];
            M.gmn2[e] = prm.g * std::pow(acc / 3, 2);
        }
    }
    if (prm.wetting_drying) {  // :213-250
        std::vector<double> Tbl((size_t)nd * 3, 0.0), Tlb(3 * (size_t)nd, 0.0);
        const double bl[3][3] = {{1.0, 1.0, 1.0}, {-1.0, -1.0, 2.0}, {-1.0, 1.0, 0.0}};  // T_bl rows 0..2
        for (int d = 0; d < 3; ++d)
            for (int v = 0; v < 3; ++v) Tbl[(size_t)d * 3 + v] = bl[d][v];
        const double lb[3][3] = {{1.0 / 3.0, -1.0 / 6.0, -1.0 / 2.0},
                                 {1.0 / 3.0, -1.0 / 6.0, 1.0 / 2.0},
                                 {1.0 / 3.0, 1.0 / 3.0, 0.0}};  // T_lb columns 0..2
        for (int v = 0; v < 3; ++v)
            for (int d = 0; d < 3; ++d) Tlb[(size_t)v * nd + d] = lb[v][d];
        for (int e = 0; e < ne; ++e) {
            std::vector<double> q(3 * (size_t)nd);
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < nd; ++d) q[(size_t)k * nd + d] = M.q0[(size_t)(k * nd + d) * ne + e];
            std::vector<double> qv = matmul(q, 3, nd, Tbl, 3);
            bool set_wet = true;
            for (int v = 0; v < 3; ++v) {
                const double h = qv[v] + tri[e].z[v];
                if (h <= M.h_o + M.h_thr) {
                    qv[v] = M.h_o - tri[e].z[v];
                    set_wet = false;
                }
            }
            if (!set_wet) {
                M.wet0[e] = 0;
                for (int v = 0; v < 3; ++v) qv[3 + v] = 0.0, qv[6 + v] = 0.0;
                q = matmul(qv, 3, 3, Tlb, nd);
                for (int k = 0; k < 3; ++k)
                    for (int d = 0; d < nd; ++d) M.q0[(size_t)(k * nd + d) * ne + e] = q[(size_t)k * nd + d];
            }
        }
    }
    if (prm.slope_limiting) {  // :252-326
        M.sl_median.assign(6 * (size_t)ne, 0.0);
        M.sl_alpha.assign(6 * (size_t)ne, 0.0);
        M.sl_rsq.assign(3 * (size_t)ne, 0.0);
        M.sl_len.assign(3 * (size_t)ne, 0.0);
        M.sl_radius.assign(ne, 0.0);
        M.sl_bath_bc.assign(ne, 0.0);
        std::vector<std::array<double, 2>> bc(ne);
        for (int e = 0; e < ne; ++e) {
            const Tri& T = tri[e];
            bc[e] = {(T.x[0] + T.x[1] + T.x[2]) / 3.0, (T.y[0] + T.y[1] + T.y[2]) / 3.0};
        }
        for (int e = 0; e < ne; ++e) {
            const Tri& T = tri[e];
            M.sl_bath_bc[e] = (T.z[0] * (1.0 / 3.0) + T.z[1] * (1.0 / 3.0)) + T.z[2] * (1.0 / 3.0);
            const double area = 0.5 * (T.x[0] * (T.y[1] - T.y[2]) + T.x[1] * (T.y[2] - T.y[0]) + T.x[2] * (T.y[0] - T.y[1]));
            double r = 1.0 / (4 * area);
            for (int s = 0; s < 3; ++s) r *= std::hypot(T.x[s] - T.x[(s + 1) % 3], T.y[s] - T.y[(s + 1) % 3]);
            M.sl_radius[e] = r;
            double mid[3][2], nbc[3][2];
            for (int k = 0; k < 3; ++k) {
                M.sl_len[(size_t)k * ne + e] =
                    std::hypot(T.x[(k + 1) % 3] - T.x[(k + 2) % 3], T.y[(k + 1) % 3] - T.y[(k + 2) % 3]);
                mid[k][0] = (T.x[(k + 1) % 3] + T.x[(k + 2) % 3]) / 2.0;
                mid[k][1] = (T.y[(k + 1) % 3] + T.y[(k + 2) % 3]) / 2.0;
                const double mx = mid[k][0] - bc[e][0], my = mid[k][1] - bc[e][1];
                const double nrm = std::sqrt(mx * mx + my * my);
                M.sl_median[(size_t)(2 * k) * ne + e] = mx / nrm;
                M.sl_median[(size_t)(2 * k + 1) * ne + e] = my / nrm;
            }
            for (int k = 0; k < 3; ++k) {
                const int n = nbr[3 * e + k];
                if (n >= 0) {
                    nbc[k][0] = bc[n][0], nbc[k][1] = bc[n][1];
                } else {  // reflected barycentre across a boundary face
                    nbc[k][0] = 2.0 * mid[k][0] - bc[e][0];
                    nbc[k][1] = 2.0 * mid[k][1] - bc[e][1];
                }
            }
            for (int b = 0; b < 3; ++b) {
                const int e1 = b, e2 = (b + 1) % 3;
                const double A00 = nbc[e1][0] - bc[e][0], A10 = nbc[e1][1] - bc[e][1];
                const double A01 = nbc[e2][0] - bc[e][0], A11 = nbc[e2][1] - bc[e][1];
                const double b0 = mid[b][0] - bc[e][0], b1 = mid[b][1] - bc[e][1];
                const double det = A00 * A11 - A10 * A01;
                const double invdet = 1.0 / det;
                const double I00 = A11 * invdet, I10 = -A10 * invdet, I01 = -A01 * invdet, I11 = A00 * invdet;
                M.sl_alpha[(size_t)(2 * b) * ne + e] = I00 * b0 + I01 * b1;
                M.sl_alpha[(size_t)(2 * b + 1) * ne + e] = I10 * b0 + I11 * b1;
                M.sl_rsq[(size_t)b * ne + e] = b0 * b0 + b1 * b1;
            }
        }
    }
    return std::string();
}

// ---------------------------------------------------------------------------------------
namespace {

// Appends columns of an element-fastest [ncol][ne] array with padding to the merged [ncol][ne_tot]
// array; `map` transforms the values (e.g. element/edge index offsets).
template <typename T, typename F>
void append_cols(std::vector<T>& out, const std::vector<T>& in, int ncol, int ne_in, int ne_tot, int base,
                 T pad, F map) {
    if (out.empty()) out.assign((size_t)ncol * ne_tot, pad);
    for (int c = 0; c < ncol; ++c)
        for (int e = 0; e < ne_in; ++e) out[(size_t)c * ne_tot + base + e] = map(in[(size_t)c * ne_in + e]);
}
template <typename T>
void append_cols(std::vector<T>& out, const std::vector<T>& in, int ncol, int ne_in, int ne_tot, int base, T pad) {
    append_cols(out, in, ncol, ne_in, ne_tot, base, pad, [](T v) { return v; });
}

}  // namespace

std::string merge_models(const std::vector<const Model*>& parts, Model& out) {
    if (parts.empty()) return "empty ensemble";
    const Model& a = *parts[0];
    for (const Model* p : parts) {
        const Model& b = *p;
        const rkdg_params &x = a.prm, &y = b.prm;
        if (b.p != a.p || b.stepper.nstages != a.stepper.nstages || b.stepper.order != a.stepper.order)
            return "ensemble instances of one group need equal p and Runge-Kutta scheme";
        if (x.g != y.g || x.friction != y.friction || x.friction_cf != y.friction_cf ||
            x.wetting_drying != y.wetting_drying || b.h_o != a.h_o || x.slope_limiting != y.slope_limiting ||
            b.sl_M != a.sl_M || b.sl_nu != a.sl_nu || x.problem != y.problem || x.function_source != y.function_source ||
            b.reference_face_order != a.reference_face_order)
            return "ensemble instances need equal physics options (g, friction, wetting-drying, limiter, problem)";
        if (!b.sp_vol.empty()) return "no spherical correction in ensembles";
        if (b.ncon != a.ncon || b.freq != a.freq || b.ffact != a.ffact || b.eqarg != a.eqarg)
            return "ensemble instances need one harmonic constituent list";
        for (int i = 0; i < 4; ++i)
            if (x.problem_param[i] != y.problem_param[i]) return "ensemble instances need equal problem parameters";
    }
    if (a.stepper.nstages < 2) return "ensembles need a Runge-Kutta scheme of at least 2 stages";
    out = Model();
    out.p = a.p, out.ndof = a.ndof, out.ngp = a.ngp, out.ngpe = a.ngpe;
    out.master = a.master, out.prm = a.prm, out.h_o = a.h_o, out.h_thr = a.h_thr, out.sl_M = a.sl_M, out.sl_nu = a.sl_nu;
    out.reference_face_order = a.reference_face_order;
    out.ncon = a.ncon, out.freq = a.freq, out.ffact = a.ffact, out.eqarg = a.eqarg;
    out.stepper = a.stepper;
    auto& E = out.ens;
    E.n = (int)parts.size();
    int ne_tot = 0;
    for (const Model* p : parts) {
        E.base.push_back(ne_tot);
        E.ne.push_back(p->ne);
        E.dt.push_back(p->stepper.dt);
        E.ramp_duration.push_back(p->stepper.ramp_duration);
        ne_tot += (p->ne + kEnsembleTile - 1) / kEnsembleTile * kEnsembleTile;
        if ((size_t)ne_tot * 3 * 6 * sizeof(double) >= ((size_t)1 << 32))
            return "ensemble too large: state arrays exceed 4 GiB of 32-bit buffer offsets";
        out.closed_forms |= p->closed_forms;
    }
    out.ne = ne_tot;
    E.tile_inst.assign(ne_tot / 64, 0);
    for (int i = 0; i < E.n; ++i) {
        const int t0 = E.base[i] / 64, t1 = (i + 1 < E.n ? E.base[i + 1] : ne_tot) / 64;
        for (int t = t0; t < t1; ++t) E.tile_inst[t] = i;
    }
    const int nd = a.ndof;
    int nedge_tot = 0;
    for (const Model* p : parts) nedge_tot += p->nedge;
    out.egeom.assign(3 * (size_t)nedge_tot, 0.0);
    int edge_base = 0, force_base = 0, levee_base = 0;
    for (int i = 0; i < E.n; ++i) {
        const Model& b = *parts[i];
        const int base = E.base[i], ne = b.ne;
        append_cols(out.geom, b.geom, 5, ne, ne_tot, base, 0.0);
        append_cols(out.bath, b.bath, nd, ne, ne_tot, base, 0.0);
        append_cols(out.vxy, b.vxy, 6, ne, ne_tot, base, 0.0);
        append_cols(out.bath_vrtx, b.bath_vrtx, 3, ne, ne_tot, base, 0.0);
        append_cols(out.gmn2, b.gmn2, 1, ne, ne_tot, base, 0.0);
        append_cols(out.nb, b.nb, 3, ne, ne_tot, base, -1, [base](int v) { return v >= 0 ? v + base : v; });
        append_cols(out.edge, b.edge, 3, ne, ne_tot, base, 0, [edge_base](int v) { return v + edge_base; });
        append_cols(out.meta, b.meta, 1, ne, ne_tot, base, 0u);
        append_cols(out.neighbor, b.neighbor, 3, ne, ne_tot, base, -1, [base](int v) { return v >= 0 ? v + base : v; });
        if (a.prm.slope_limiting) {
            append_cols(out.sl_median, b.sl_median, 6, ne, ne_tot, base, 0.0);
            append_cols(out.sl_alpha, b.sl_alpha, 6, ne, ne_tot, base, 0.0);
            append_cols(out.sl_rsq, b.sl_rsq, 3, ne, ne_tot, base, 0.0);
            append_cols(out.sl_len, b.sl_len, 3, ne, ne_tot, base, 0.0);
            append_cols(out.sl_radius, b.sl_radius, 1, ne, ne_tot, base, 0.0);
            append_cols(out.sl_bath_bc, b.sl_bath_bc, 1, ne, ne_tot, base, 0.0);
            append_cols(out.sl_face_slot, b.sl_face_slot, 3, ne, ne_tot, base, 0);
        }
        append_cols(out.q0, b.q0, 3 * nd, ne, ne_tot, base, 0.0);
        append_cols(out.wet0, b.wet0, 1, ne, ne_tot, base, (uint8_t)0);
        // edge records: interfaces first (their aux record is a levee record), then boundaries
        // (their aux record is a tide/flow forcing record)
        const int ned = b.nedge;
        for (int c = 0; c < 3; ++c)
            std::copy(b.egeom.begin() + (size_t)c * ned, b.egeom.begin() + (size_t)(c + 1) * ned,
                      out.egeom.begin() + (size_t)c * nedge_tot + edge_base);
        for (int k = 0; k < ned; ++k) {
            const int r = b.eforce[k];
            out.eforce.push_back(r < 0 ? -1 : r + (k < b.nintf ? levee_base : force_base));
        }
        out.famp.insert(out.famp.end(), b.famp.begin(), b.famp.end());
        out.fphase.insert(out.fphase.end(), b.fphase.begin(), b.fphase.end());
        out.levee_gp.insert(out.levee_gp.end(), b.levee_gp.begin(), b.levee_gp.end());
        edge_base += ned;
        if (b.ncon) force_base += (int)(b.famp.size() / ((size_t)b.ncon * b.ngpe));
        levee_base += (int)(b.levee_gp.size() / (3 * (size_t)b.ngpe));
    }
    out.nedge = edge_base;
    out.nintf = 0;  // mixed; the merged eforce already carries the right record bases
    return "";
}

double l2_residual(const Model& M, const std::vector<double>& q, double t) {
    const Rule2 rule = dunavant_rule(20);
    const int ngp = (int)rule.w.size(), nd = M.ndof, ne = M.ne;
    std::vector<double> phi((size_t)nd * ngp);
    dubiner_eval(M.p, rule.z, phi.data(), nullptr, nullptr);
    const std::vector<double> psi = shape_psi(rule.z);
    double res = 0;
    for (int e = 0; e < ne; ++e) {
        double x[3], y[3];
        for (int k = 0; k < 3; ++k) x[k] = M.vxy[(size_t)k * ne + e], y[k] = M.vxy[(size_t)(3 + k) * ne + e];
        const double adet = M.geom[4 * (size_t)ne + e];
        double l2v[3] = {0, 0, 0};
        for (int g = 0; g < ngp; ++g) {
            const double gx = x[0] * psi[g] + x[1] * psi[ngp + g] + x[2] * psi[2 * ngp + g];
            const double gy = y[0] * psi[g] + y[1] * psi[ngp + g] + y[2] * psi[2 * ngp + g];
            double tv[3];
            problem_ic_host(M.prm, t, gx, gy, tv);
            for (int k = 0; k < 3; ++k) {
                double u = q[(size_t)(k * nd) * ne + e] * phi[g];
                for (int d = 1; d < nd; ++d) u = u + q[(size_t)(k * nd + d) * ne + e] * phi[(size_t)d * ngp + g];
                const double diff = tv[k] - u;
                const double term = (diff * diff) * rule.w[g];
                l2v[k] = g == 0 ? term : l2v[k] + term;
            }
        }
        double L2 = 0;
        for (int k = 0; k < 3; ++k) L2 += l2v[k] * adet;
        res += L2;
    }
    return std::sqrt(res);
}

}  // namespace rkdg
