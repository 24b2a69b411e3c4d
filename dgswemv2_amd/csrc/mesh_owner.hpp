// Owner of the arrays a library-built rkdg_mesh_desc points into (generator and ADCIRC reader).
#pragma once

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/rkdg.h"

namespace rkdg {
void set_create_error(const std::string& msg);  // the message rkdg_create_error() returns

// Runs the body of an extern "C" host entry point: no exception crosses the C-ABI; any that
// escapes the body becomes RKDG_ERR_ARG with its message in rkdg_create_error().
template <typename F>
rkdg_status guarded(const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        set_create_error(std::string(what) + ": out of host memory");
    } catch (const std::exception& ex) {
        set_create_error(std::string(what) + ": " + ex.what());
    } catch (...) {
        set_create_error(std::string(what) + ": unknown error");
    }
    return RKDG_ERR_ARG;
}
}  // namespace rkdg

struct rkdg_mesh {
    std::vector<double> xyz;
    std::vector<int32_t> elt;
    std::vector<int32_t> bedge;
    std::vector<uint8_t> btype;
    std::vector<double> freq, ff, eq, amp, phase, manning;
    std::vector<int32_t> levee_nodes;
    std::vector<double> levee_param;
    std::vector<int32_t> global_elt;  // submeshes: global element index of each local element

    void fill(rkdg_mesh_desc* d) const {
        d->n_nodes = (int32_t)(xyz.size() / 3);
        d->n_elements = (int32_t)(elt.size() / 3);
        d->node_xyz = xyz.data();
        d->elt_nodes = elt.data();
        d->n_bound_edges = (int32_t)btype.size();
        d->bound_edge_nodes = bedge.empty() ? nullptr : bedge.data();
        d->bound_edge_type = btype.empty() ? nullptr : btype.data();
        d->n_constituents = (int32_t)freq.size();
        d->frequency = freq.empty() ? nullptr : freq.data();
        d->forcing_fact = ff.empty() ? nullptr : ff.data();
        d->equilib_arg = eq.empty() ? nullptr : eq.data();
        d->node_amplitude = amp.empty() ? nullptr : amp.data();
        d->node_phase = phase.empty() ? nullptr : phase.data();
        d->node_manning_n = manning.empty() ? nullptr : manning.data();
        d->n_levee_pairs = (int32_t)(levee_nodes.size() / 2);
        d->levee_nodes = levee_nodes.empty() ? nullptr : levee_nodes.data();
        d->levee_param = levee_param.empty() ? nullptr : levee_param.data();
    }
};
