// Mesh partitioning for the distributed-boundary path: recursive coordinate bisection of the
// elements and the submesh of one part (its elements plus ghost copies of the neighbours it
// gathers from other parts).
//
// The reference partitions with METIS into submeshes whose cut edges become distributed
// boundaries (partitioner/partition.cpp, preprocessor/mesh_metadata.cpp:139-223), and exchanges
// edge traces per stage (rkdg_swe_proc_ompi_step.hpp:38-102).  Here a submesh keeps its ghost
// elements as ordinary elements in global element order: an element of the part then sees exactly
// the operands it sees in the whole mesh (neighbour modes, which side of an interface is the
// lower ID, the order of its faces), so a partitioned run equals the whole-mesh run bit for bit
// once the ghosts hold their owners' current state before each stage.  Levee faces are interfaces
// between the elements either side of the barrier (rkdg_swe_dbc_distributed_levee.hpp carries them
// across parts in the reference), so the element across a levee is a ghost like any other
// neighbour.  Faces of a ghost that lead out of the submesh (levee faces included) are typed land;
// ghosts' own results are overwritten by the exchange.
#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rkdg.h"
#include "mesh_owner.hpp"

namespace {

struct PartError {
    std::string msg;
};

uint64_t ekey(long a, long b) { return (uint64_t)std::min(a, b) << 32 | (uint64_t)std::max(a, b); }

// Recursive coordinate bisection: split `idx` (element indices) across parts [p0, p0 + np) in
// proportion to the part counts, along the longer extent of the barycentres (ties by index).
void rcb(std::vector<int>& idx, size_t lo, size_t hi, int p0, int np, const std::vector<double>& cx,
         const std::vector<double>& cy, int32_t* part) {
    if (np == 1) {
        for (size_t i = lo; i < hi; ++i) part[idx[i]] = p0;
        return;
    }
    double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
    for (size_t i = lo; i < hi; ++i) {
        x0 = std::min(x0, cx[idx[i]]), x1 = std::max(x1, cx[idx[i]]);
        y0 = std::min(y0, cy[idx[i]]), y1 = std::max(y1, cy[idx[i]]);
    }
    const std::vector<double>& c = (x1 - x0 >= y1 - y0) ? cx : cy;
    const int n1 = np / 2;
    const size_t mid = lo + (size_t)((double)(hi - lo) * n1 / np + 0.5);
    std::nth_element(idx.begin() + lo, idx.begin() + mid, idx.begin() + hi,
                     [&](int a, int b) { return c[a] < c[b] || (c[a] == c[b] && a < b); });
    rcb(idx, lo, mid, p0, n1, cx, cy, part);
    rcb(idx, mid, hi, p0 + n1, np - n1, cx, cy, part);
}

}  // namespace

extern "C" rkdg_status rkdg_partition_rcb(const rkdg_mesh_desc* d, int32_t nparts, int32_t* part) {
    return rkdg::guarded("rkdg_partition_rcb", [&]() -> rkdg_status {
    if (!d || !part || nparts < 1 || d->n_elements < nparts || !d->node_xyz || !d->elt_nodes) {
        rkdg::set_create_error("bad arguments");
        return RKDG_ERR_ARG;
    }
    const int ne = d->n_elements;
    std::vector<double> cx(ne), cy(ne);
    for (int e = 0; e < ne; ++e) {
        double x = 0.0, y = 0.0;
        for (int k = 0; k < 3; ++k) {
            const int v = d->elt_nodes[3 * e + k];
            if (v < 0 || v >= d->n_nodes) {
                rkdg::set_create_error("element node out of range");
                return RKDG_ERR_ARG;
            }
            x += d->node_xyz[3 * v], y += d->node_xyz[3 * v + 1];
        }
        cx[e] = x / 3.0, cy[e] = y / 3.0;
    }
    std::vector<int> idx(ne);
    std::iota(idx.begin(), idx.end(), 0);
    rcb(idx, 0, ne, 0, nparts, cx, cy, part);
    return RKDG_OK;
    });
}

extern "C" rkdg_status rkdg_submesh(const rkdg_mesh_desc* d, const int32_t* part, int32_t p, rkdg_mesh** out,
                                    rkdg_mesh_desc* sub, const int32_t** global_element) {
    if (!d || !part || !out || !sub) {
        rkdg::set_create_error("null argument");
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    try {
        const long ne = d->n_elements, nn = d->n_nodes, nc = d->n_constituents;
        std::unordered_map<uint64_t, std::array<long, 2>> owners;  // edge -> its elements (-1: none)
        owners.reserve(3 * (size_t)ne);
        for (long e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                auto it = owners.emplace(ekey(d->elt_nodes[3 * e + (k + 1) % 3], d->elt_nodes[3 * e + (k + 2) % 3]),
                                         std::array<long, 2>{e, -1});
                if (!it.second) it.first->second[1] = e;
            }
        std::unordered_map<uint64_t, int> btype;
        for (long i = 0; i < d->n_bound_edges; ++i)
            btype[ekey(d->bound_edge_nodes[2 * i], d->bound_edge_nodes[2 * i + 1])] = d->bound_edge_type[i];
        // A levee face's neighbour is the element across the barrier: the one edge whose nodes are
        // the barrier partners of the face's nodes (the levee interface of the whole mesh).
        std::unordered_map<long, long> partner;
        for (long i = 0; i < d->n_levee_pairs; ++i) {
            const long f = d->levee_nodes[2 * i], b = d->levee_nodes[2 * i + 1];
            if (f < 0 || f >= nn || b < 0 || b >= nn) throw PartError{"levee node out of range"};
            partner[f] = b, partner[b] = f;
        }
        auto across = [&](long e, int k) -> long {  // neighbour of face k of e, levees included
            const long a = d->elt_nodes[3 * e + (k + 1) % 3], b = d->elt_nodes[3 * e + (k + 2) % 3];
            const auto& o = owners[ekey(a, b)];
            if (o[1] >= 0) return o[0] == e ? o[1] : o[0];
            auto t = btype.find(ekey(a, b));
            if (t == btype.end() || t->second != RKDG_BC_LEVEE) return -1;
            auto pa = partner.find(a), pb = partner.find(b);
            if (pa == partner.end() || pb == partner.end()) throw PartError{"levee edge without barrier node pairs"};
            auto x = owners.find(ekey(pa->second, pb->second));
            if (x == owners.end() || x->second[1] >= 0) throw PartError{"levee edge without a single partner element"};
            return x->second[0];
        };
        // local elements: the part's own and their neighbours, in global order
        std::vector<char> local(ne, 0);
        long nown = 0;
        for (long e = 0; e < ne; ++e) {
            if (part[e] != p) continue;
            ++nown, local[e] = 1;
            for (int k = 0; k < 3; ++k) {
                const long nb = across(e, k);
                if (nb >= 0) local[nb] = 1;
            }
        }
        if (nown == 0) throw PartError{"part " + std::to_string(p) + " has no elements"};
        std::unique_ptr<rkdg_mesh> M(new rkdg_mesh());
        std::vector<long> nloc(nn, -1);
        std::vector<char> nused(nn, 0);
        for (long e = 0; e < ne; ++e)
            if (local[e]) {
                M->global_elt.push_back((int32_t)e);
                for (int k = 0; k < 3; ++k) nused[d->elt_nodes[3 * e + k]] = 1;
            }
        long nl = 0;
        for (long v = 0; v < nn; ++v)
            if (nused[v]) {
                nloc[v] = nl++;
                for (int k = 0; k < 3; ++k) M->xyz.push_back(d->node_xyz[3 * v + k]);
                if (d->node_manning_n) M->manning.push_back(d->node_manning_n[v]);
            }
        for (int32_t e : M->global_elt)
            for (int k = 0; k < 3; ++k) M->elt.push_back((int32_t)nloc[d->elt_nodes[3 * e + k]]);
        // boundary edges: the global ones among the local elements' faces (levee faces whose
        // partner is local stay levees), and land on ghost faces that lead out of the submesh
        for (int32_t e : M->global_elt)
            for (int k = 0; k < 3; ++k) {
                const long a = d->elt_nodes[3 * e + (k + 1) % 3], b = d->elt_nodes[3 * e + (k + 2) % 3];
                const bool interior = owners[ekey(a, b)][1] >= 0;
                const long nb = across(e, k);
                int t;
                if (nb >= 0 && !local[nb]) {
                    t = RKDG_BC_LAND;  // only ghosts have neighbours outside the submesh
                } else if (!interior) {
                    auto it = btype.find(ekey(a, b));
                    if (it == btype.end()) throw PartError{"boundary edge without a code"};
                    t = it->second;
                } else {
                    continue;
                }
                M->bedge.push_back((int32_t)nloc[a]), M->bedge.push_back((int32_t)nloc[b]);
                M->btype.push_back((uint8_t)t);
            }
        // barrier node pairs whose nodes are both in the submesh (those of the levee faces kept)
        for (long i = 0; i < d->n_levee_pairs; ++i) {
            const long f = d->levee_nodes[2 * i], b = d->levee_nodes[2 * i + 1];
            if (!nused[f] || !nused[b]) continue;
            M->levee_nodes.push_back((int32_t)nloc[f]), M->levee_nodes.push_back((int32_t)nloc[b]);
            for (int c = 0; c < 3; ++c) M->levee_param.push_back(d->levee_param[3 * i + c]);
        }
        if (nc > 0) {
            M->freq.assign(d->frequency, d->frequency + nc);
            M->ff.assign(d->forcing_fact, d->forcing_fact + nc);
            M->eq.assign(d->equilib_arg, d->equilib_arg + nc);
            for (long v = 0; v < nn; ++v)
                if (nused[v])
                    for (long c = 0; c < nc; ++c) {
                        M->amp.push_back(d->node_amplitude ? d->node_amplitude[v * nc + c] : 0.0);
                        M->phase.push_back(d->node_phase ? d->node_phase[v * nc + c] : 0.0);
                    }
        }
        M->fill(sub);
        if (global_element) *global_element = M->global_elt.data();
        *out = M.release();
        return RKDG_OK;
    } catch (const PartError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("partition error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}

// ---- locality renumbering ----------------------------------------------------------------
// Elements in the order of a Hilbert curve through their barycentres (2^16 x 2^16 grid over the
// mesh's bounding box, ties by index); nodes in order of first use by the renumbered elements.
// Neighbours then sit close in memory, so the stage kernel's neighbour gathers hit in L2 on meshes
// whose file order is arbitrary.  The reference's partitioner plays this role for its submeshes.

namespace {

uint64_t hilbert_d(uint32_t n, uint32_t x, uint32_t y) {  // position along the curve of order log2(n)
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {  // rotate the quadrant
            if (rx == 1) x = s - 1 - x, y = s - 1 - y;
            std::swap(x, y);
        }
    }
    return d;
}

}  // namespace

extern "C" rkdg_status rkdg_mesh_renumber(const rkdg_mesh_desc* d, rkdg_mesh** out, rkdg_mesh_desc* sub,
                                          const int32_t** old_element) {
    if (!d || !out || !sub || !d->node_xyz || !d->elt_nodes || d->n_elements <= 0) {
        rkdg::set_create_error("bad arguments");
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    try {
        const long ne = d->n_elements, nn = d->n_nodes, nc = d->n_constituents;
        std::vector<double> cx(ne), cy(ne);
        double x0 = 1e300, x1 = -1e300, y0 = 1e300, y1 = -1e300;
        for (long e = 0; e < ne; ++e) {
            double x = 0.0, y = 0.0;
            for (int k = 0; k < 3; ++k) {
                const int v = d->elt_nodes[3 * e + k];
                if (v < 0 || v >= nn) throw PartError{"element node out of range"};
                x += d->node_xyz[3 * v], y += d->node_xyz[3 * v + 1];
            }
            cx[e] = x / 3.0, cy[e] = y / 3.0;
            x0 = std::min(x0, cx[e]), x1 = std::max(x1, cx[e]), y0 = std::min(y0, cy[e]), y1 = std::max(y1, cy[e]);
        }
        const uint32_t n = 1u << 16;
        const double sx = x1 > x0 ? (n - 1) / (x1 - x0) : 0.0, sy = y1 > y0 ? (n - 1) / (y1 - y0) : 0.0;
        std::vector<std::pair<uint64_t, long>> key(ne);
        for (long e = 0; e < ne; ++e)
            key[e] = {hilbert_d(n, (uint32_t)((cx[e] - x0) * sx), (uint32_t)((cy[e] - y0) * sy)), e};
        std::sort(key.begin(), key.end());
        std::unique_ptr<rkdg_mesh> M(new rkdg_mesh());
        std::vector<long> nnew(nn, -1);
        long nl = 0;
        for (const auto& kv : key) {
            const long e = kv.second;
            M->global_elt.push_back((int32_t)e);
            for (int k = 0; k < 3; ++k) {
                const int v = d->elt_nodes[3 * e + k];
                if (nnew[v] < 0) nnew[v] = nl++;
                M->elt.push_back((int32_t)nnew[v]);
            }
        }
        for (long v = 0; v < nn; ++v)
            if (nnew[v] < 0) nnew[v] = nl++;  // nodes no element uses keep a place at the end
        std::vector<long> nold(nn);
        for (long v = 0; v < nn; ++v) nold[nnew[v]] = v;
        for (long i = 0; i < nn; ++i) {
            const long v = nold[i];
            for (int k = 0; k < 3; ++k) M->xyz.push_back(d->node_xyz[3 * v + k]);
            if (d->node_manning_n) M->manning.push_back(d->node_manning_n[v]);
            for (long c = 0; c < nc; ++c) {
                M->amp.push_back(d->node_amplitude ? d->node_amplitude[v * nc + c] : 0.0);
                M->phase.push_back(d->node_phase ? d->node_phase[v * nc + c] : 0.0);
            }
        }
        if (nc > 0) {
            M->freq.assign(d->frequency, d->frequency + nc);
            M->ff.assign(d->forcing_fact, d->forcing_fact + nc);
            M->eq.assign(d->equilib_arg, d->equilib_arg + nc);
        }
        for (long i = 0; i < d->n_bound_edges; ++i) {
            M->bedge.push_back((int32_t)nnew[d->bound_edge_nodes[2 * i]]);
            M->bedge.push_back((int32_t)nnew[d->bound_edge_nodes[2 * i + 1]]);
            M->btype.push_back(d->bound_edge_type[i]);
        }
        for (long i = 0; i < d->n_levee_pairs; ++i) {
            M->levee_nodes.push_back((int32_t)nnew[d->levee_nodes[2 * i]]);
            M->levee_nodes.push_back((int32_t)nnew[d->levee_nodes[2 * i + 1]]);
            for (int k = 0; k < 3; ++k) M->levee_param.push_back(d->levee_param[3 * i + k]);
        }
        M->fill(sub);
        if (old_element) *old_element = M->global_elt.data();
        *out = M.release();
        return RKDG_OK;
    } catch (const PartError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("partition error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}
