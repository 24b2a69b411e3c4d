// ADCIRC mesh (.14) and boundary-data (.bcis) reader producing an rkdg_mesh_desc.
//
// Follows the reference's preprocessor: AdcircFormat (preprocessor/ADCIRC_reader/
// adcirc_format.cpp:3-130) for the file layout, AdcircFormat::get_ibtype (:214-258) for the
// boundary code of every mesh edge with one element, get_internal_node_pair (:260-310) for the
// partner of a barrier node, and SWE::Inputs::read_bcis (problem_input/swe_inputs.cpp:212-312)
// for tide/flow constituents and levee coefficients.
#include <algorithm>
#include <array>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <new>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rkdg.h"
#include "mesh_owner.hpp"

namespace {

struct Segment {          // one land/flow/barrier/function/outflow boundary string
    int ibtype = 0;
    std::vector<long> nodes;  // NBVV
    std::vector<long> back;   // IBCONN (barriers)
    std::vector<std::array<double, 3>> barrier;  // BARINTH, BARINCFSB, BARINCFSP
};

struct Fort14 {
    std::string name;
    std::map<long, std::array<double, 3>> nodes;
    std::map<long, std::array<long, 3>> elements;
    std::vector<std::vector<long>> open;  // NBDV
    std::vector<Segment> seg;             // NBOU segments
    std::vector<std::vector<long>> generic;
};

struct ReadError {
    std::string msg;
};

// Reads the next line and returns it as a stream (throws at end of file).
std::istringstream next_line(std::istream& is, const char* what) {
    std::string line;
    if (!std::getline(is, line)) throw ReadError{std::string("unexpected end of file while reading ") + what};
    return std::istringstream(line);
}

// A count read from a line of its own (an ADCIRC count field): missing, unparsable or negative
// counts are file errors, not sizes.
long read_count(std::istringstream ss, const char* what) {
    long n = -1;
    if (!(ss >> n) || n < 0) throw ReadError{std::string("bad ") + what + " count"};
    return n;
}

Fort14 read_fort14(const std::string& path) {
    std::ifstream ifs(path);
    if (!ifs) throw ReadError{"ADCIRC mesh file " + path + " was not found"};
    Fort14 F;
    std::getline(ifs, F.name);
    long ne = 0, nn = 0;
    {
        auto ss = next_line(ifs, "element/node counts");
        if (!(ss >> ne >> nn) || ne < 0 || nn < 0) throw ReadError{"bad element/node counts"};
    }
    if (ne <= 0 || nn <= 0) throw ReadError{"bad element/node counts"};
    for (long i = 0; i < nn; ++i) {
        long id;
        std::array<double, 3> v;
        auto ss = next_line(ifs, "nodes");
        if (!(ss >> id >> v[0] >> v[1] >> v[2])) throw ReadError{"malformed node line"};
        if (id < 0) throw ReadError{"node ID is negative"};
        if (!F.nodes.emplace(id, v).second) throw ReadError{"node defined twice"};
    }
    for (long i = 0; i < ne; ++i) {
        long id, k;
        std::array<long, 3> v;
        auto ss = next_line(ifs, "elements");
        if (!(ss >> id >> k >> v[0] >> v[1] >> v[2])) throw ReadError{"malformed element line"};
        if (id < 0) throw ReadError{"element ID is negative"};
        if (k != 3) throw ReadError{"only triangles are supported"};
        if (!F.elements.emplace(id, v).second) throw ReadError{"element defined twice"};
    }
    long n;
    const long nope = read_count(next_line(ifs, "open boundaries"), "open boundary");
    next_line(ifs, "open boundary nodes");
    for (long b = 0; b < nope; ++b) {
        n = read_count(next_line(ifs, "open boundary"), "open boundary node");
        std::vector<long> v(n);
        for (long i = 0; i < n; ++i) next_line(ifs, "open boundary node") >> v[i];
        F.open.push_back(std::move(v));
    }
    const long nbou = read_count(next_line(ifs, "land boundaries"), "land boundary");
    next_line(ifs, "land boundary nodes");
    for (long b = 0; b < nbou; ++b) {
        Segment S;
        {
            auto ss = next_line(ifs, "boundary segment");
            if (!(ss >> n >> S.ibtype) || n < 0) throw ReadError{"bad boundary segment node count"};
        }
        const int t = S.ibtype % 10;
        if (t == 0 || t == 1 || t == 2 || S.ibtype == 77 || S.ibtype == 88) {
            S.nodes.resize(n);
            for (long i = 0; i < n; ++i) next_line(ifs, "boundary node") >> S.nodes[i];
        } else if (t == 4) {
            S.nodes.resize(n), S.back.resize(n), S.barrier.resize(n);
            for (long i = 0; i < n; ++i) {
                auto ss = next_line(ifs, "barrier node");
                if (!(ss >> S.nodes[i] >> S.back[i] >> S.barrier[i][0] >> S.barrier[i][1] >> S.barrier[i][2]))
                    throw ReadError{"malformed barrier line"};
            }
        } else {
            throw ReadError{"undefined boundary type in ADCIRC mesh: " + std::to_string(S.ibtype)};
        }
        F.seg.push_back(std::move(S));
    }
    std::string line;
    if (std::getline(ifs, line) && line.find_first_not_of(" \t\r") != std::string::npos) {  // generic boundaries
        const long ngen = read_count(std::istringstream(line), "generic boundary");
        next_line(ifs, "generic boundary nodes");
        for (long b = 0; b < ngen; ++b) {
            n = read_count(next_line(ifs, "generic boundary"), "generic boundary node");
            std::vector<long> v(n);
            for (long i = 0; i < n; ++i) next_line(ifs, "generic boundary node") >> v[i];
            F.generic.push_back(std::move(v));
        }
    }
    return F;
}

// a and b consecutive in the string, either order (AdcircFormat::has_edge, :312-336)
bool has_edge(const std::vector<long>& s, long a, long b) {
    for (size_t i = 0; i + 1 < s.size(); ++i)
        if ((s[i] == a && s[i + 1] == b) || (s[i] == b && s[i + 1] == a)) return true;
    return false;
}

// AdcircFormat::is_levee_end (:338-356): the edge joining the first (or last) front and back nodes
bool is_levee_end(const Fort14& F, long a, long b) {
    const long lo = std::min(a, b), hi = std::max(a, b);
    for (const Segment& S : F.seg) {
        if (S.ibtype % 10 != 4 || S.nodes.empty()) continue;
        const long f0 = S.nodes.front(), b0 = S.back.front(), f1 = S.nodes.back(), b1 = S.back.back();
        if ((lo == std::min(f0, b0) && hi == std::max(f0, b0)) || (lo == std::min(f1, b1) && hi == std::max(f1, b1)))
            return true;
    }
    return false;
}

// AdcircFormat::get_ibtype (:214-258).  Island strings (IBTYPE % 10 == 1) are read but, as in
// the reference, not matched here.
int ibtype_of(const Fort14& F, long a, long b) {
    for (const auto& s : F.open)
        if (has_edge(s, a, b)) return RKDG_BC_TIDE;
    for (const Segment& S : F.seg) {
        const int t = S.ibtype % 10;
        if (t == 0 && has_edge(S.nodes, a, b)) return RKDG_BC_LAND;
        if (t == 2 && has_edge(S.nodes, a, b)) return RKDG_BC_FLOW;
        if (t == 4 && (has_edge(S.nodes, a, b) || has_edge(S.back, a, b))) return RKDG_BC_LEVEE;
        if (S.ibtype == 77 && has_edge(S.nodes, a, b)) return RKDG_BC_FUNCTION;
        if (S.ibtype == 88 && has_edge(S.nodes, a, b)) return 4;  // outflow: rejected by the engine
    }
    for (const auto& s : F.generic)
        if (has_edge(s, a, b)) return RKDG_BC_LAND;
    if (is_levee_end(F, a, b)) return RKDG_BC_LAND;
    throw ReadError{"boundary not found for node pair (" + std::to_string(a) + ", " + std::to_string(b) + ")"};
}

// SWE::Inputs::read_bcis (problem_input/swe_inputs.cpp:212-312): tide / flow blocks (constituents
// with amplitude and phase per node) and the levee block (coefficients per front/back node pair).
// `node` maps a file node ID to its index.
struct Bcis {
    std::vector<std::array<double, 3>> cons;                       // frequency, forcing factor, eq. argument
    std::map<long, std::vector<std::array<double, 2>>> forcing;    // node -> (amplitude, phase) per constituent
    std::map<std::pair<long, long>, std::array<double, 3>> levee;  // (front, back) -> H, C_sub, C_super
    std::vector<std::pair<long, long>> levee_order;                // pairs in file order

    void apply(rkdg_mesh& M, long nn) const {
        if (cons.empty()) return;
        const size_t nc = cons.size();
        for (auto& c : cons) M.freq.push_back(c[0]), M.ff.push_back(c[1]), M.eq.push_back(c[2]);
        M.amp.assign((size_t)nn * nc, 0.0);
        M.phase.assign((size_t)nn * nc, 0.0);
        for (auto& kv : forcing) {
            if (kv.first < 0 || kv.first >= nn) throw ReadError{"forcing node out of range"};
            for (size_t c = 0; c < nc; ++c)
                M.amp[(size_t)kv.first * nc + c] = kv.second[c][0], M.phase[(size_t)kv.first * nc + c] = kv.second[c][1];
        }
    }
};

template <typename NodeFn>
Bcis read_bcis(const char* path, NodeFn&& node) {
    std::ifstream ifs(path);
    if (!ifs) throw ReadError{std::string("BC/ISP data file ") + path + " was not found"};
    Bcis B;
    std::string line;
    while (std::getline(ifs, line)) {
        long btype = -1, nnodes = 0;
        std::istringstream(line) >> btype >> nnodes;
        if (btype == RKDG_BC_TIDE || btype == RKDG_BC_FLOW) {
            if (nnodes < 0) throw ReadError{"bad forcing node count"};
            const long ncon = read_count(next_line(ifs, "constituent count"), "constituent");
            std::vector<std::array<double, 3>> here;
            std::map<long, std::vector<std::array<double, 2>>> block;
            for (long c = 0; c < ncon; ++c) {
                std::array<double, 3> cc;
                next_line(ifs, "constituent") >> cc[0] >> cc[1] >> cc[2];
                here.push_back(cc);
                for (long i = 0; i < nnodes; ++i) {
                    long nd;
                    double amp, ph;
                    auto ss = next_line(ifs, "forcing node");
                    if (!(ss >> nd >> amp >> ph)) throw ReadError{"malformed forcing line"};
                    block[node(nd)].push_back({amp, ph});
                }
            }
            if (B.cons.empty())
                B.cons = here;
            else if (B.cons != here)
                throw ReadError{"tide/flow blocks with different constituent lists are not supported"};
            for (auto& kv : block) {
                auto it = B.forcing.find(kv.first);
                if (it != B.forcing.end() && it->second != kv.second)
                    throw ReadError{"node " + std::to_string(kv.first) + " forced by two different blocks"};
                B.forcing[kv.first] = kv.second;
            }
        } else if (btype == RKDG_BC_LEVEE) {
            for (long i = 0; i < nnodes; ++i) {
                long f, b;
                std::array<double, 3> v;
                auto ss = next_line(ifs, "levee node pair");
                if (!(ss >> f >> b >> v[0] >> v[1] >> v[2])) throw ReadError{"malformed levee line"};
                const std::pair<long, long> key{node(f), node(b)};
                if (!B.levee.count(key)) B.levee_order.push_back(key);
                B.levee[key] = v;
            }
        }
    }
    return B;
}

}  // namespace

extern "C" rkdg_status rkdg_mesh_read_adcirc(const char* fort14_path, const char* bcis_path, rkdg_mesh** out,
                                             rkdg_mesh_desc* desc) {
    if (!fort14_path || !out || !desc) {
        rkdg::set_create_error("null argument");
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    try {
        Fort14 F = read_fort14(fort14_path);
        const long nn = (long)F.nodes.size(), ne = (long)F.elements.size();
        // Node and element IDs become their rank among the file's IDs.  Every ordering the
        // reference derives from IDs (element container order, interface in = lower ID, edge
        // keys by (min, max) node) is preserved by this monotone relabelling.
        std::unordered_map<long, long> nrank;
        nrank.reserve(2 * (size_t)nn);
        std::unique_ptr<rkdg_mesh> M(new rkdg_mesh());
        M->xyz.resize(3 * (size_t)nn);
        for (const auto& kv : F.nodes) {
            const long r = (long)nrank.size();
            nrank[kv.first] = r;
            for (int k = 0; k < 3; ++k) M->xyz[3 * (size_t)r + k] = kv.second[k];
        }
        auto node = [&](long name) {
            auto it = nrank.find(name);
            if (it == nrank.end()) throw ReadError{"reference to undefined node " + std::to_string(name)};
            return it->second;
        };
        M->elt.resize(3 * (size_t)ne);
        long er = 0;
        for (const auto& kv : F.elements) {
            for (int k = 0; k < 3; ++k) M->elt[3 * (size_t)er + k] = (int32_t)node(kv.second[k]);
            ++er;
        }
        for (auto& s : F.open)
            for (long& v : s) v = node(v);
        for (auto& S : F.seg) {
            for (long& v : S.nodes) v = node(v);
            for (long& v : S.back) v = node(v);
        }
        for (auto& s : F.generic)
            for (long& v : s) v = node(v);
        // edges with one element (MeshMetaData, mesh_metadata.cpp:31-85) and their boundary code
        std::unordered_map<uint64_t, int> count;
        count.reserve(3 * (size_t)ne);
        for (long e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                const uint64_t a = (uint64_t)M->elt[3 * e + (k + 1) % 3], b = (uint64_t)M->elt[3 * e + (k + 2) % 3];
                ++count[std::min(a, b) << 32 | std::max(a, b)];
            }
        for (long e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k) {
                const long a = M->elt[3 * e + (k + 1) % 3], b = M->elt[3 * e + (k + 2) % 3];
                const uint64_t key = (uint64_t)std::min(a, b) << 32 | (uint64_t)std::max(a, b);
                if (count[key] != 1) continue;
                M->bedge.push_back((int32_t)a), M->bedge.push_back((int32_t)b);
                M->btype.push_back((uint8_t)ibtype_of(F, a, b));
            }

        // .bcis: tide/flow constituents per node, levee coefficients per node pair
        Bcis B;
        if (bcis_path) B = read_bcis(bcis_path, node);
        B.apply(*M, nn);
        const auto& levee_bcis = B.levee;
        // barrier node pairs (front NBVV, back IBCONN); coefficients from the .bcis levee block
        // (what the reference uses, swe_pre_create_intface.hpp:50-69), else from the .14 barrier
        // columns BARINTH / BARINCFSB / BARINCFSP
        for (const Segment& S : F.seg) {
            if (S.ibtype % 10 != 4) continue;
            for (size_t i = 0; i < S.nodes.size(); ++i) {
                std::array<double, 3> v = S.barrier[i];
                if (bcis_path) {
                    auto it = levee_bcis.find({S.nodes[i], S.back[i]});
                    if (it == levee_bcis.end()) it = levee_bcis.find({S.back[i], S.nodes[i]});
                    if (it == levee_bcis.end())
                        throw ReadError{"unable to find levee data for node pair (" + std::to_string(S.nodes[i]) + ", " +
                                        std::to_string(S.back[i]) + ")"};
                    v = it->second;
                }
                M->levee_nodes.push_back((int32_t)S.nodes[i]), M->levee_nodes.push_back((int32_t)S.back[i]);
                for (int k = 0; k < 3; ++k) M->levee_param.push_back(v[k]);
            }
        }
        M->fill(desc);
        *out = M.release();
        return RKDG_OK;
    } catch (const ReadError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("mesh file error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}

namespace {

// Splits a set of edges into node strings whose consecutive pairs are exactly those edges, each
// used once: walks start at nodes of odd degree (string ends), then what is left is closed loops
// (all degrees even), each written with its first node repeated at the end.  Deterministic: starts
// in order of first appearance in the edge list.
std::vector<std::vector<long>> chain_edges(const std::vector<std::array<long, 2>>& E) {
    std::unordered_map<long, std::vector<size_t>> adj;
    std::vector<long> order;  // nodes by first appearance
    for (size_t i = 0; i < E.size(); ++i)
        for (long v : E[i]) {
            auto& a = adj[v];
            if (a.empty()) order.push_back(v);
            a.push_back(i);
        }
    std::vector<char> used(E.size(), 0);
    std::unordered_map<long, size_t> cursor;
    auto walk = [&](long start) {
        std::vector<long> s{start};
        long cur = start;
        for (;;) {
            auto& a = adj[cur];
            size_t& c = cursor[cur];
            while (c < a.size() && used[a[c]]) ++c;
            if (c == a.size()) break;
            const size_t i = a[c];
            used[i] = 1;
            cur = E[i][0] == cur ? E[i][1] : E[i][0];
            s.push_back(cur);
        }
        return s;
    };
    std::vector<std::vector<long>> out;
    for (long v : order)
        if (adj[v].size() % 2 == 1)
            while (true) {
                auto s = walk(v);
                if (s.size() < 2) break;
                out.push_back(std::move(s));
            }
    for (size_t i = 0; i < E.size(); ++i)
        if (!used[i]) out.push_back(walk(E[i][0]));
    return out;
}

struct WriteError {
    std::string msg;
};

}  // namespace

extern "C" rkdg_status rkdg_mesh_write_adcirc(const rkdg_mesh_desc* d, const char* name, const char* fort14_path,
                                              const char* bcis_path) {
    if (!d || !fort14_path) {
        rkdg::set_create_error("null argument");
        return RKDG_ERR_ARG;
    }
    try {
        const long nn = d->n_nodes, ne = d->n_elements, nb = d->n_bound_edges, nl = d->n_levee_pairs;
        const int nc = d->n_constituents;
        if (nn <= 0 || ne <= 0 || !d->node_xyz || !d->elt_nodes) throw WriteError{"empty mesh"};
        if (nb < 0 || (nb > 0 && (!d->bound_edge_nodes || !d->bound_edge_type))) throw WriteError{"bad boundary edges"};
        if (nl < 0 || (nl > 0 && (!d->levee_nodes || !d->levee_param))) throw WriteError{"bad levee pairs"};
        auto check_node = [&](long v) {
            if (v < 0 || v >= nn) throw WriteError{"node index " + std::to_string(v) + " out of range"};
            return v;
        };
        for (long i = 0; i < 3 * ne; ++i) check_node(d->elt_nodes[i]);
        for (long i = 0; i < 2 * nb; ++i) check_node(d->bound_edge_nodes[i]);
        for (long i = 0; i < 2 * nl; ++i) check_node(d->levee_nodes[i]);

        // boundary edges by code; the writer's string kinds (AdcircFormat::write_to,
        // adcirc_format.cpp:151-212): open boundaries (NBDV) for tide, IBTYPE 0 land, 2 flow,
        // 77 function, 88 outflow, 4 internal barrier (NBVV front, IBCONN back)
        std::map<int, std::vector<std::array<long, 2>>> by_type;
        std::map<std::pair<long, long>, int> edge_type;
        for (long i = 0; i < nb; ++i) {
            const long a = d->bound_edge_nodes[2 * i], b = d->bound_edge_nodes[2 * i + 1];
            const int t = d->bound_edge_type[i];
            if (t != RKDG_BC_LAND && t != RKDG_BC_TIDE && t != RKDG_BC_FLOW && t != RKDG_BC_FUNCTION &&
                t != RKDG_BC_LEVEE && t != 4)
                throw WriteError{"boundary code " + std::to_string(t) + " has no ADCIRC boundary type"};
            if (t != RKDG_BC_LEVEE) by_type[t].push_back({a, b});
            edge_type[{std::min(a, b), std::max(a, b)}] = t;
        }
        auto is_levee_edge = [&](long a, long b) {
            auto it = edge_type.find({std::min(a, b), std::max(a, b)});
            return it != edge_type.end() && it->second == RKDG_BC_LEVEE;
        };
        // barrier segments: the levee pairs in order, split where the front or back nodes of
        // consecutive pairs are not joined by a levee edge
        std::vector<std::vector<long>> bar;  // pair indices
        for (long i = 0; i < nl; ++i) {
            const int32_t* L = d->levee_nodes;
            if (i == 0 || !is_levee_edge(L[2 * i - 2], L[2 * i]) || !is_levee_edge(L[2 * i - 1], L[2 * i + 1]))
                bar.emplace_back();
            bar.back().push_back(i);
        }
        // every levee edge must be a consecutive front or back pair of a barrier segment
        std::map<std::pair<long, long>, int> covered;
        for (const auto& s : bar)
            for (size_t k = 1; k < s.size(); ++k)
                for (int side = 0; side < 2; ++side) {
                    const long a = d->levee_nodes[2 * s[k - 1] + side], b = d->levee_nodes[2 * s[k] + side];
                    covered[{std::min(a, b), std::max(a, b)}] = 1;
                }
        for (const auto& kv : edge_type)
            if (kv.second == RKDG_BC_LEVEE && !covered.count(kv.first))
                throw WriteError{"levee edge (" + std::to_string(kv.first.first) + ", " +
                                 std::to_string(kv.first.second) + ") joins no two consecutive levee pairs"};

        FILE* f = std::fopen(fort14_path, "w");
        if (!f) throw WriteError{std::string("cannot open ") + fort14_path};
        std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
        std::fprintf(f, "%s\n%ld %ld\n", name && *name ? name : "mesh", ne, nn);
        for (long i = 0; i < nn; ++i)
            std::fprintf(f, "%ld %.17g %.17g %.17g\n", i, d->node_xyz[3 * i], d->node_xyz[3 * i + 1], d->node_xyz[3 * i + 2]);
        for (long e = 0; e < ne; ++e)
            std::fprintf(f, "%ld 3 %d %d %d\n", e, d->elt_nodes[3 * e], d->elt_nodes[3 * e + 1], d->elt_nodes[3 * e + 2]);
        const auto open = chain_edges(by_type[RKDG_BC_TIDE]);
        size_t neta = 0;
        for (const auto& s : open) neta += s.size();
        std::fprintf(f, "%zu = Number of open boundaries\n%zu = Total number of open boundary nodes\n", open.size(), neta);
        for (size_t k = 0; k < open.size(); ++k) {
            std::fprintf(f, "%zu = Number of nodes for open boundary %zu\n", open[k].size(), k + 1);
            for (long v : open[k]) std::fprintf(f, "%ld\n", v);
        }
        std::vector<std::pair<int, std::vector<long>>> land;  // (IBTYPE, nodes)
        const std::pair<int, int> kinds[] = {{RKDG_BC_LAND, 0}, {RKDG_BC_FLOW, 2}, {RKDG_BC_FUNCTION, 77}, {4, 88}};
        for (const auto& k : kinds)
            for (auto& s : chain_edges(by_type[k.first])) land.emplace_back(k.second, std::move(s));
        size_t nvel = 0;
        for (const auto& s : land) nvel += s.second.size();
        for (const auto& s : bar) nvel += 2 * s.size();
        std::fprintf(f, "%zu = Number of land boundaries\n%zu = Total number of land boundary nodes\n",
                     land.size() + bar.size(), nvel);
        size_t seg = 0;
        for (const auto& s : land) {
            std::fprintf(f, "%zu %d = Number of nodes for land boundary %zu\n", s.second.size(), s.first, ++seg);
            for (long v : s.second) std::fprintf(f, "%ld\n", v);
        }
        for (const auto& s : bar) {
            std::fprintf(f, "%zu 4 = Number of nodes for land boundary %zu\n", s.size(), ++seg);
            for (long i : s)
                std::fprintf(f, "%d %d %.17g %.17g %.17g\n", d->levee_nodes[2 * i], d->levee_nodes[2 * i + 1],
                             d->levee_param[3 * i], d->levee_param[3 * i + 1], d->levee_param[3 * i + 2]);
        }
        std::fprintf(f, "0 = Number of generic boundaries\n0 = Total number of generic boundary nodes\n");
        if (std::ferror(f)) throw WriteError{std::string("write error on ") + fort14_path};
        guard.reset();

        if (bcis_path) {
            FILE* g = std::fopen(bcis_path, "w");
            if (!g) throw WriteError{std::string("cannot open ") + bcis_path};
            std::unique_ptr<FILE, int (*)(FILE*)> gguard(g, &std::fclose);
            // tide / flow blocks (SWE::Inputs::read_bcis, swe_inputs.cpp:212-296): the nodes of
            // that boundary kind with their amplitude and phase per constituent
            if (nc > 0) {
                if (!d->frequency || !d->forcing_fact || !d->equilib_arg || !d->node_amplitude || !d->node_phase)
                    throw WriteError{"constituents without node amplitudes/phases"};
                for (int t : {(int)RKDG_BC_TIDE, (int)RKDG_BC_FLOW}) {
                    std::vector<long> nodes;
                    for (const auto& e : by_type[t]) nodes.push_back(e[0]), nodes.push_back(e[1]);
                    std::sort(nodes.begin(), nodes.end());
                    nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
                    if (nodes.empty()) continue;
                    std::fprintf(g, "%d %zu = %s boundary\n%d = number of constituents\n", t, nodes.size(),
                                 t == RKDG_BC_TIDE ? "tide" : "flow", nc);
                    for (int c = 0; c < nc; ++c) {
                        std::fprintf(g, "%.17g %.17g %.17g = frequency, forcing factor, equilibrium argument\n",
                                     d->frequency[c], d->forcing_fact[c], d->equilib_arg[c]);
                        for (long v : nodes)
                            std::fprintf(g, "%ld %.17g %.17g\n", v, d->node_amplitude[(size_t)v * nc + c],
                                         d->node_phase[(size_t)v * nc + c]);
                    }
                }
            }
            // levee block (:298-310): front, back, barrier height, sub-/supercritical coefficients
            if (nl > 0) {
                std::fprintf(g, "%d %ld = levee boundary\n", (int)RKDG_BC_LEVEE, nl);
                for (long i = 0; i < nl; ++i)
                    std::fprintf(g, "%d %d %.17g %.17g %.17g\n", d->levee_nodes[2 * i], d->levee_nodes[2 * i + 1],
                                 d->levee_param[3 * i], d->levee_param[3 * i + 1], d->levee_param[3 * i + 2]);
            }
            if (std::ferror(g)) throw WriteError{std::string("write error on ") + bcis_path};
        }
        return RKDG_OK;
    } catch (const WriteError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("mesh file error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}

// ---- MeshMetaData (.meta) files -------------------------------------------------------------
// Layout of MeshMetaData::write_to / MeshMetaData(const std::string&)
// (preprocessor/mesh_metadata.cpp:87-137, element and node records mesh_metadata.hpp:11-72): the
// mesh name, "<n> = number of elements", one line per element "ID nfaces node0 node1 node2
// neighbour0 neighbour1 neighbour2 type0 type1 type2" with face k opposite node k, the neighbour
// 4294967295 on a boundary face (the element across the barrier on a levee face) and each type
// one raw byte (SWE::BoundaryTypes: 0 land, 1 tide, 2 flow, 3 function, 4 outflow, 128 interior,
// 129 levee), then "<n> = number of nodes" and one line per node "ID x y z".

namespace {

constexpr unsigned kNoNeighbor = 4294967295u;  // DEFAULT_ID (general_definitions.hpp:230)
constexpr int kInternal = 128;                 // INTERNAL (utilities/edge_types.hpp:4)

uint64_t edge_key(long a, long b) { return (uint64_t)std::min(a, b) << 32 | (uint64_t)std::max(a, b); }

}  // namespace

extern "C" rkdg_status rkdg_mesh_write_meta(const rkdg_mesh_desc* d, const char* name, const char* path) {
    if (!d || !path) {
        rkdg::set_create_error("null argument");
        return RKDG_ERR_ARG;
    }
    try {
        const long nn = d->n_nodes, ne = d->n_elements, nb = d->n_bound_edges, nl = d->n_levee_pairs;
        if (nn <= 0 || ne <= 0 || !d->node_xyz || !d->elt_nodes) throw WriteError{"empty mesh"};
        if (nb < 0 || (nb > 0 && (!d->bound_edge_nodes || !d->bound_edge_type))) throw WriteError{"bad boundary edges"};
        if (nl < 0 || (nl > 0 && !d->levee_nodes)) throw WriteError{"bad levee pairs"};
        for (long i = 0; i < 3 * ne; ++i)
            if (d->elt_nodes[i] < 0 || d->elt_nodes[i] >= nn) throw WriteError{"element node out of range"};
        // faces by edge: the (up to two) elements holding it
        std::unordered_map<uint64_t, std::vector<std::pair<long, int>>> faces;
        faces.reserve(3 * (size_t)ne);
        for (long e = 0; e < ne; ++e)
            for (int k = 0; k < 3; ++k)
                faces[edge_key(d->elt_nodes[3 * e + (k + 1) % 3], d->elt_nodes[3 * e + (k + 2) % 3])].push_back({e, k});
        std::unordered_map<uint64_t, int> btype;
        for (long i = 0; i < nb; ++i)
            btype[edge_key(d->bound_edge_nodes[2 * i], d->bound_edge_nodes[2 * i + 1])] = d->bound_edge_type[i];
        // levee partner of a node: front <-> back (AdcircFormat::get_internal_node_pair)
        std::unordered_map<long, long> partner;
        for (long i = 0; i < nl; ++i) {
            partner[d->levee_nodes[2 * i]] = d->levee_nodes[2 * i + 1];
            partner[d->levee_nodes[2 * i + 1]] = d->levee_nodes[2 * i];
        }
        FILE* f = std::fopen(path, "wb");
        if (!f) throw WriteError{std::string("cannot open ") + path};
        std::unique_ptr<FILE, int (*)(FILE*)> guard(f, &std::fclose);
        std::fprintf(f, "%s\n%ld = number of elements\n", name && *name ? name : "mesh", ne);
        for (long e = 0; e < ne; ++e) {
            unsigned nbr[3];
            int type[3];
            for (int k = 0; k < 3; ++k) {
                const long a = d->elt_nodes[3 * e + (k + 1) % 3], b = d->elt_nodes[3 * e + (k + 2) % 3];
                const auto& fl = faces[edge_key(a, b)];
                if (fl.size() == 2) {
                    nbr[k] = (unsigned)(fl[0].first == e ? fl[1].first : fl[0].first), type[k] = kInternal;
                    continue;
                }
                if (fl.size() != 1) throw WriteError{"edge shared by more than two elements"};
                auto it = btype.find(edge_key(a, b));
                if (it == btype.end())
                    throw WriteError{"boundary face (" + std::to_string(a) + ", " + std::to_string(b) + ") has no code"};
                type[k] = it->second, nbr[k] = kNoNeighbor;
                if (type[k] == RKDG_BC_LEVEE) {
                    auto pa = partner.find(a), pb = partner.find(b);
                    if (pa == partner.end() || pb == partner.end())
                        throw WriteError{"levee face (" + std::to_string(a) + ", " + std::to_string(b) +
                                         ") has no levee node pairs"};
                    auto fo = faces.find(edge_key(pa->second, pb->second));
                    if (fo == faces.end() || fo->second.size() != 1)
                        throw WriteError{"levee face (" + std::to_string(a) + ", " + std::to_string(b) +
                                         ") has no face across the barrier"};
                    nbr[k] = (unsigned)fo->second[0].first;
                }
            }
            std::fprintf(f, "%ld 3 %d %d %d %u %u %u %c %c %c\n", e, d->elt_nodes[3 * e], d->elt_nodes[3 * e + 1],
                         d->elt_nodes[3 * e + 2], nbr[0], nbr[1], nbr[2], (char)type[0], (char)type[1], (char)type[2]);
        }
        std::fprintf(f, "%ld = number of nodes\n", nn);
        for (long i = 0; i < nn; ++i)
            std::fprintf(f, "%ld %.17g %.17g %.17g\n", i, d->node_xyz[3 * i], d->node_xyz[3 * i + 1], d->node_xyz[3 * i + 2]);
        if (std::ferror(f)) throw WriteError{std::string("write error on ") + path};
        return RKDG_OK;
    } catch (const WriteError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("mesh file error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}

extern "C" rkdg_status rkdg_mesh_read_meta(const char* meta_path, const char* bcis_path, rkdg_mesh** out,
                                           rkdg_mesh_desc* desc) {
    if (!meta_path || !out || !desc) {
        rkdg::set_create_error("null argument");
        return RKDG_ERR_ARG;
    }
    *out = nullptr;
    try {
        std::ifstream ifs(meta_path, std::ios::binary);
        if (!ifs) throw ReadError{std::string("meta mesh file ") + meta_path + " was not found"};
        std::string name;
        ifs >> name;
        ifs.ignore(1 << 20, '\n');
        long ne = 0, nn = 0;
        ifs >> ne;
        ifs.ignore(1 << 20, '\n');
        if (!ifs || ne <= 0) throw ReadError{"bad element count"};
        struct Elt {
            std::array<long, 3> node;
            std::array<unsigned char, 3> type;
        };
        std::map<long, Elt> elts;
        for (long i = 0; i < ne; ++i) {
            long id, nf;
            Elt E;
            unsigned long nbr;
            ifs >> id >> nf;
            if (!ifs || nf != 3) throw ReadError{"malformed element record (only triangles are supported)"};
            for (int k = 0; k < 3; ++k) ifs >> E.node[k];
            for (int k = 0; k < 3; ++k) ifs >> nbr;
            for (int k = 0; k < 3; ++k) {
                char c;
                ifs >> c;  // one raw byte per face (skips whitespace, as operator>> on unsigned char does)
                E.type[k] = (unsigned char)c;
            }
            ifs.ignore(1 << 20, '\n');
            if (!ifs) throw ReadError{"malformed element record"};
            if (!elts.emplace(id, E).second) throw ReadError{"element defined twice"};
        }
        ifs >> nn;
        ifs.ignore(1 << 20, '\n');
        if (!ifs || nn <= 0) throw ReadError{"bad node count"};
        std::map<long, std::array<double, 3>> nodes;
        for (long i = 0; i < nn; ++i) {
            long id;
            std::array<double, 3> v;
            ifs >> id >> v[0] >> v[1] >> v[2];
            ifs.ignore(1 << 20, '\n');
            if (!ifs) throw ReadError{"malformed node record"};
            if (!nodes.emplace(id, v).second) throw ReadError{"node defined twice"};
        }
        // IDs become their rank among the file's IDs (order-preserving, as for the .14)
        std::unordered_map<long, long> nrank;
        std::unique_ptr<rkdg_mesh> M(new rkdg_mesh());
        M->xyz.resize(3 * (size_t)nn);
        for (const auto& kv : nodes) {
            const long r = (long)nrank.size();
            nrank[kv.first] = r;
            for (int k = 0; k < 3; ++k) M->xyz[3 * (size_t)r + k] = kv.second[k];
        }
        auto node = [&](long id) {
            auto it = nrank.find(id);
            if (it == nrank.end()) throw ReadError{"reference to undefined node " + std::to_string(id)};
            return it->second;
        };
        M->elt.reserve(3 * (size_t)ne);
        for (const auto& kv : elts)
            for (int k = 0; k < 3; ++k) M->elt.push_back((int32_t)node(kv.second.node[k]));
        // boundary faces in element, then face order (the order the .14 reader lists them in)
        long e = 0;
        for (const auto& kv : elts) {
            for (int k = 0; k < 3; ++k) {
                const int t = kv.second.type[k];
                if (t == kInternal) continue;
                if (t > RKDG_BC_LEVEE || (t > 4 && t < kInternal))
                    throw ReadError{"undefined boundary type " + std::to_string(t) + " in meta mesh"};
                M->bedge.push_back(M->elt[3 * e + (k + 1) % 3]), M->bedge.push_back(M->elt[3 * e + (k + 2) % 3]);
                M->btype.push_back((uint8_t)t);
            }
            ++e;
        }
        Bcis B;
        if (bcis_path) B = read_bcis(bcis_path, node);
        B.apply(*M, nn);
        // levee pairs and coefficients: the .bcis levee block (the .meta holds no node pairs)
        for (const auto& key : B.levee_order) {
            const auto& v = B.levee.at(key);
            M->levee_nodes.push_back((int32_t)key.first), M->levee_nodes.push_back((int32_t)key.second);
            for (int k = 0; k < 3; ++k) M->levee_param.push_back(v[k]);
        }
        if (M->levee_nodes.empty())
            for (uint8_t t : M->btype)
                if (t == RKDG_BC_LEVEE) throw ReadError{"levee faces need the .bcis levee block"};
        M->fill(desc);
        *out = M.release();
        return RKDG_OK;
    } catch (const ReadError& err) {
        rkdg::set_create_error(err.msg);
    } catch (const std::bad_alloc&) {
        rkdg::set_create_error("out of host memory");
    } catch (const std::exception& ex) {
        rkdg::set_create_error(std::string("mesh file error: ") + ex.what());
    }
    return RKDG_ERR_ARG;
}
