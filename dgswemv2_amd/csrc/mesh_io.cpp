// ADCIRC mesh (.14) and boundary-data (.bcis) reader producing an rkdg_mesh_desc.
//
// Follows the reference's preprocessor: AdcircFormat (preprocessor/ADCIRC_reader/
// adcirc_format.cpp:3-130) for the file layout, AdcircFormat::get_ibtype (:214-258) for the
// boundary code of every mesh edge with one element, get_internal_node_pair (:260-310) for the
// partner of a barrier node, and SWE::Inputs::read_bcis (problem_input/swe_inputs.cpp:212-312)
// for tide/flow constituents and levee coefficients.
#include <algorithm>
#include <array>
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

Fort14 read_fort14(const std::string& path) {
    std::ifstream ifs(path);
    if (!ifs) throw ReadError{"ADCIRC mesh file " + path + " was not found"};
    Fort14 F;
    std::getline(ifs, F.name);
    long ne = 0, nn = 0;
    next_line(ifs, "element/node counts") >> ne >> nn;
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
    long nope = 0, n;
    next_line(ifs, "open boundaries") >> nope;
    next_line(ifs, "open boundary nodes");
    for (long b = 0; b < nope; ++b) {
        next_line(ifs, "open boundary") >> n;
        std::vector<long> v(n);
        for (long i = 0; i < n; ++i) next_line(ifs, "open boundary node") >> v[i];
        F.open.push_back(std::move(v));
    }
    long nbou = 0;
    next_line(ifs, "land boundaries") >> nbou;
    next_line(ifs, "land boundary nodes");
    for (long b = 0; b < nbou; ++b) {
        Segment S;
        next_line(ifs, "boundary segment") >> n >> S.ibtype;
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
        long ngen = 0;
        std::istringstream(line) >> ngen;
        next_line(ifs, "generic boundary nodes");
        for (long b = 0; b < ngen; ++b) {
            next_line(ifs, "generic boundary") >> n;
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
        std::map<std::pair<long, long>, std::array<double, 3>> levee_bcis;
        std::vector<std::array<double, 3>> cons;
        std::map<long, std::vector<std::array<double, 2>>> node_forcing;  // node -> (amplitude, phase) per constituent
        if (bcis_path) {
            std::ifstream ifs(bcis_path);
            if (!ifs) throw ReadError{std::string("BC/ISP data file ") + bcis_path + " was not found"};
            std::string line;
            while (std::getline(ifs, line)) {
                long btype = -1, nnodes = 0;
                std::istringstream(line) >> btype >> nnodes;
                if (btype == RKDG_BC_TIDE || btype == RKDG_BC_FLOW) {
                    long ncon = 0;
                    next_line(ifs, "constituent count") >> ncon;
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
                    if (cons.empty())
                        cons = here;
                    else if (cons != here)
                        throw ReadError{"tide/flow blocks with different constituent lists are not supported"};
                    for (auto& kv : block) {
                        auto it = node_forcing.find(kv.first);
                        if (it != node_forcing.end() && it->second != kv.second)
                            throw ReadError{"node " + std::to_string(kv.first) + " forced by two different blocks"};
                        node_forcing[kv.first] = kv.second;
                    }
                } else if (btype == RKDG_BC_LEVEE) {
                    for (long i = 0; i < nnodes; ++i) {
                        long f, b;
                        std::array<double, 3> v;
                        auto ss = next_line(ifs, "levee node pair");
                        if (!(ss >> f >> b >> v[0] >> v[1] >> v[2])) throw ReadError{"malformed levee line"};
                        levee_bcis[{node(f), node(b)}] = v;
                    }
                }
            }
        }
        if (!cons.empty()) {
            const size_t nc = cons.size();
            for (auto& c : cons) M->freq.push_back(c[0]), M->ff.push_back(c[1]), M->eq.push_back(c[2]);
            M->amp.assign((size_t)nn * nc, 0.0);
            M->phase.assign((size_t)nn * nc, 0.0);
            for (auto& kv : node_forcing) {
                if (kv.first < 0 || kv.first >= nn) throw ReadError{"forcing node out of range"};
                for (size_t c = 0; c < nc; ++c)
                    M->amp[(size_t)kv.first * nc + c] = kv.second[c][0], M->phase[(size_t)kv.first * nc + c] = kv.second[c][1];
            }
        }
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
    }
    return RKDG_ERR_ARG;
}
