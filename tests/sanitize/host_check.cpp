// Host-side sanitizer check, CPU only (built by tests/sanitize/Makefile with g++ and
// -fsanitize=address,undefined; run by tests/test_host_sanitizers.py).
//
// Exercises the host code of librkdg that needs no GPU -- the quad-mesh generator, the ADCIRC
// .14 / .bcis and .meta readers and writers (with malformed files), the coordinate-bisection
// partitioner, submeshes, Hilbert renumbering, the model builder (geometry, tables, ensemble
// merge), the L2 residual and the modal / VTU writers -- and the FP64 oracle (whole stages with
// wetting-drying, the limiter, levees, tide / flow forcing; the single-element limiter), so that
// AddressSanitizer and UndefinedBehaviorSanitizer see every array access they make.
// Usage: host_check <weir.14> <weir.bcis> <sample_fort.14> <scratch dir>.  Exit 0 when every check
// holds (a sanitizer report aborts the process with a nonzero status).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/rkdg.h"
#include "model.hpp"
#include "writers.hpp"

namespace rkdg {
// the library defines this next to rkdg_create (api.cpp, a HIP translation unit): the last
// message of a call without a context
static std::string g_err;
void set_create_error(const std::string& msg) { g_err = msg; }
}  // namespace rkdg

extern "C" {
void* orc_create(const rkdg_mesh_desc* md, const rkdg_params* prm);
const char* orc_error();
void orc_destroy(void* h);
int orc_step(void* h, int nsteps);
void orc_get_state(void* h, double* q, uint8_t* wet, uint8_t* limited, uint8_t* troubled);
void orc_set_state(void* h, const double* q, const uint8_t* wet);
double orc_l2_residual(void* h);
int orc_apply_limiter(void* h);
void orc_cs_limit_element(double g, double M, double nu, const double* q_lin, const double* q_nb, double bath_bar,
                          const double* median, const double* alpha, const double* r_sq, double* q_vrtx, double* delta,
                          int* limited);
}

static int g_fail = 0;
#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            ++g_fail;                                                         \
        }                                                                     \
    } while (0)

static rkdg_params params(int p, int ns, int order, double dt) {
    rkdg_params prm;
    std::memset(&prm, 0, sizeof prm);
    prm.p = p, prm.nstages = ns, prm.order = order, prm.dt = dt, prm.g = 9.81, prm.ic_constant = 1;
    prm.h_o = 0.1, prm.sl_M = 1.0e-8, prm.sl_nu = 1.5, prm.sph_R = 6378200.0;
    return prm;
}

static rkdg_quad_mesh_spec spec(int m, int pattern, int bathy, double jitter) {
    rkdg_quad_mesh_spec s;
    std::memset(&s, 0, sizeof s);
    s.x1 = 0, s.y1 = 0, s.x2 = 1000.0 * m, s.y2 = 1000.0 * m, s.m = m, s.n = m + 1, s.pattern = pattern;
    s.bc[0] = RKDG_BC_LAND, s.bc[1] = RKDG_BC_FLOW, s.bc[2] = RKDG_BC_LAND, s.bc[3] = RKDG_BC_TIDE;
    s.bathy_kind = bathy, s.b0 = 10.0, s.alpha = 1e-7, s.bmin = 5.0, s.bmax = 400.0, s.n_modes = 16;
    s.jitter = jitter, s.seed = 3, s.step_x = 500.0 * m, s.step_b_left = 10.0, s.step_b_right = 8.5;
    s.frequency = 1.405189e-4, s.forcing_fact = 1.0, s.amplitude = 0.5, s.manning_n = 0.025;
    return s;
}

// Oracle whole stages on a mesh, then the state must be finite.
static void run_oracle(const rkdg_mesh_desc& d, const rkdg_params& prm, int steps) {
    void* h = orc_create(&d, &prm);
    if (!h) {
        std::fprintf(stderr, "oracle create failed: %s\n", orc_error());
        ++g_fail;
        return;
    }
    CHECK(orc_step(h, steps) == RKDG_OK);
    const int nd = (prm.p + 1) * (prm.p + 2) / 2;
    std::vector<double> q(3 * (size_t)nd * d.n_elements);
    std::vector<uint8_t> wet(d.n_elements), lim(3 * (size_t)d.n_elements), tr(d.n_elements);
    orc_get_state(h, q.data(), wet.data(), lim.data(), tr.data());
    size_t bad = 0;
    for (double v : q) bad += !std::isfinite(v);
    if (bad) std::fprintf(stderr, "%zu non-finite values (p %d, wd %d, problem %d)\n", bad, prm.p, prm.wetting_drying,
                          prm.problem);
    CHECK(bad == 0);
    if (prm.slope_limiting) CHECK(orc_apply_limiter(h) == RKDG_OK);
    (void)orc_l2_residual(h);
    orc_destroy(h);
}

// Model builder, L2 residual and both writers on a mesh.
static void run_model(const rkdg_mesh_desc& d, const rkdg_params& prm, const std::string& prefix) {
    for (bool ref_order : {true, false}) {
        rkdg::Model M;
        const std::string err = rkdg::build_model(d, prm, M, ref_order);
        CHECK(err.empty());
        if (!err.empty()) {
            std::fprintf(stderr, "build_model: %s\n", err.c_str());
            continue;
        }
        std::vector<double> q(3 * (size_t)M.ndof * M.ne, 0.25);
        std::vector<uint8_t> wet(M.ne, 1);
        (void)rkdg::l2_residual(M, q, 0.5);
        std::string werr;
        CHECK(rkdg::write_modal(M, q.data(), 0.0, 0, prefix, werr));
        CHECK(rkdg::write_vtu(M, q.data(), wet.data(), 0, prefix, werr));
    }
}

// Writes a mesh as .14/.bcis and .meta, reads both back, partitions and renumbers it.
static void round_trips(const rkdg_mesh_desc& d, const std::string& dir, const std::string& tag) {
    const std::string f14 = dir + "/" + tag + ".14", fb = dir + "/" + tag + ".bcis", fm = dir + "/" + tag + ".meta";
    CHECK(rkdg_mesh_write_adcirc(&d, tag.c_str(), f14.c_str(), fb.c_str()) == RKDG_OK);
    rkdg_mesh* r = nullptr;
    rkdg_mesh_desc rd;
    CHECK(rkdg_mesh_read_adcirc(f14.c_str(), fb.c_str(), &r, &rd) == RKDG_OK);
    if (r) {
        CHECK(rd.n_elements == d.n_elements && rd.n_nodes == d.n_nodes && rd.n_bound_edges == d.n_bound_edges);
        rkdg_mesh_free(r);
    }
    CHECK(rkdg_mesh_write_meta(&d, tag.c_str(), fm.c_str()) == RKDG_OK);
    r = nullptr;
    CHECK(rkdg_mesh_read_meta(fm.c_str(), fb.c_str(), &r, &rd) == RKDG_OK);
    if (r) {
        CHECK(rd.n_elements == d.n_elements);
        rkdg_mesh_free(r);
    }
    const int nparts = d.n_elements >= 8 ? 4 : 1;
    std::vector<int32_t> part(d.n_elements);
    CHECK(rkdg_partition_rcb(&d, nparts, part.data()) == RKDG_OK);
    for (int p = 0; p < nparts; ++p) {
        rkdg_mesh* s = nullptr;
        rkdg_mesh_desc sd;
        const int32_t* gid = nullptr;
        CHECK(rkdg_submesh(&d, part.data(), p, &s, &sd, &gid) == RKDG_OK);
        if (s) {
            for (int e = 0; e < sd.n_elements; ++e) CHECK(gid[e] >= 0 && gid[e] < d.n_elements);
            rkdg_mesh_free(s);
        }
    }
    rkdg_mesh* h = nullptr;
    rkdg_mesh_desc hd;
    const int32_t* old = nullptr;
    CHECK(rkdg_mesh_renumber(&d, &h, &hd, &old) == RKDG_OK);
    if (h) rkdg_mesh_free(h);
}

static void malformed(const std::string& dir) {
    const char* head = "bad\n1 3\n1 0 0 5\n2 1 0 5\n3 0 1 5\n1 3 1 2 3\n";
    const char* tails[] = {
        "1 = NOPE\n-5 = NETA\n-5 = NVDLL\n0 = NBOU\n0 = NVEL\n",  // negative node count of a string
        "-1 = NOPE\n0 = NETA\n0 = NBOU\n0 = NVEL\n",              // negative open boundary count
        "0 = NOPE\n0 = NETA\n1 = NBOU\n3 = NVEL\n-3 0\n",           // negative segment count
        "0 = NOPE\n0 = NETA\n1 = NBOU\n3 = NVEL\n3 0\n1\n",         // truncated segment
        "0 = NOPE\n0 = NETA\n1 = NBOU\n2 = NVEL\n2 5\n1\n2\n",      // undefined boundary type
        "",                                                         // no boundary sections
    };
    int k = 0;
    for (const char* t : tails) {
        const std::string f = dir + "/bad" + std::to_string(k++) + ".14";
        std::ofstream(f) << head << t;
        rkdg_mesh* m = nullptr;
        rkdg_mesh_desc d;
        CHECK(rkdg_mesh_read_adcirc(f.c_str(), nullptr, &m, &d) == RKDG_ERR_ARG);
        if (m) rkdg_mesh_free(m);
    }
    rkdg_mesh* m = nullptr;
    rkdg_mesh_desc d;
    CHECK(rkdg_mesh_read_adcirc((dir + "/missing.14").c_str(), nullptr, &m, &d) == RKDG_ERR_ARG);
    std::ofstream(dir + "/bad.meta") << "x\n-4 = number of elements\n";
    CHECK(rkdg_mesh_read_meta((dir + "/bad.meta").c_str(), nullptr, &m, &d) == RKDG_ERR_ARG);
}

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: host_check weir.14 weir.bcis sample_fort.14 dir\n");
        return 2;
    }
    const std::string dir = argv[4];
    // generated meshes: every pattern and bathymetry kind, tide + flow + land sides
    int k = 0;
    for (int pattern = 0; pattern < 3; ++pattern)
        for (int bathy = 0; bathy < 4; ++bathy) {
            rkdg_quad_mesh_spec s = spec(5 + pattern, pattern, bathy, bathy == 2 ? 0.2 : 0.0);
            rkdg_mesh* g = nullptr;
            rkdg_mesh_desc d;
            CHECK(rkdg_mesh_generate(&s, &g, &d) == RKDG_OK);
            if (!g) continue;
            const std::string tag = "gen" + std::to_string(k++);
            round_trips(d, dir, tag);
            rkdg_params prm = params(bathy == 2 ? 2 : 1, bathy == 2 ? 3 : 2, bathy == 2 ? 3 : 2, 1.0);
            prm.friction = RKDG_FRICTION_MANNING;
            prm.ramp_duration = bathy == 1 ? 0.01 : 0.0;
            // bowl (depth -alpha (x^2 + y^2): dry) and step bathymetries run with wetting-drying and the limiter
            if (bathy == 1 || bathy == 3) prm.wetting_drying = 1, prm.slope_limiting = 1, prm.h_o = 0.01;
            if (bathy == 1) prm.problem = RKDG_PROBLEM_PARABOLIC_BOWL, prm.ic_constant = 0;
            run_oracle(d, prm, 3);
            run_model(d, prm, dir + "/" + tag);
            rkdg_mesh_free(g);
        }
    // the reference's fixtures: the weir (levees, flow forcing) with the limiter, p = 1 and 2
    {
        rkdg_mesh* w = nullptr;
        rkdg_mesh_desc d;
        CHECK(rkdg_mesh_read_adcirc(argv[1], argv[2], &w, &d) == RKDG_OK);
        if (w) {
            round_trips(d, dir, "weir");
            for (int p = 1; p <= 2; ++p) {
                rkdg_params prm = params(p, 2, 2, 3.0);
                prm.slope_limiting = 1;
                prm.ze0 = 0.3;
                run_oracle(d, prm, 2);
                run_model(d, prm, dir + "/weir" + std::to_string(p));
            }
            rkdg_mesh_free(w);
        }
        rkdg_mesh* s = nullptr;
        CHECK(rkdg_mesh_read_adcirc(argv[3], nullptr, &s, &d) == RKDG_OK);
        if (s) {
            round_trips(d, dir, "sample");
            rkdg_mesh_free(s);
        }
    }
    // an ensemble merge of two generated meshes of one order
    {
        rkdg_quad_mesh_spec a = spec(4, 0, 2, 0.2), b = spec(6, 1, 2, 0.2);
        rkdg_mesh *ga = nullptr, *gb = nullptr;
        rkdg_mesh_desc da, db;
        CHECK(rkdg_mesh_generate(&a, &ga, &da) == RKDG_OK && rkdg_mesh_generate(&b, &gb, &db) == RKDG_OK);
        if (ga && gb) {
            rkdg_params prm = params(2, 3, 3, 1.0);
            rkdg::Model ma, mb, merged;
            CHECK(rkdg::build_model(da, prm, ma, false).empty() && rkdg::build_model(db, prm, mb, false).empty());
            CHECK(rkdg::merge_models({&ma, &mb}, merged).empty());
            CHECK(merged.ens.n == 2 && merged.ne >= ma.ne + mb.ne);
        }
        if (ga) rkdg_mesh_free(ga);
        if (gb) rkdg_mesh_free(gb);
    }
    malformed(dir);
    // the limiting step of one element (the hand-worked case of tests/test_oracle_limiter.py)
    const double ql[9] = {0.6, 1.2, 1.2, -0.8, 0.6, 0.2, 0.0, -0.1, 0.1};
    const double qn[9] = {1.1, 0, 0, 1.3, 0.2, 0, 0.8, 0, 0};
    const double med[6] = {1, 0, 0, 1, -0.6, -0.8}, al[6] = {0.5, 0.5, 0.25, 0.75, 1.0, 0.2}, rs[3] = {1, 1, 1};
    double qv[9], dl[9];
    int lim[3];
    orc_cs_limit_element(1.0, 1e-8, 1.5, ql, qn, 3.0, med, al, rs, qv, dl, lim);
    CHECK(std::fabs(qv[0] - 0.625) < 1e-14 && lim[0] == 1);
    std::printf("host_check: %s (%d failed checks)\n", g_fail ? "FAILED" : "ok", g_fail);
    return g_fail ? 1 : 0;
}
