// Structured quad-split triangle meshes (host).
//
// Node layout, element patterns and boundary node lists follow the reference's
// quad_mesh_generator (mesh_generators/quad_mesh_generator.cpp:91-279, patterns :464-541):
// bilinear map of the unit square onto the rectangle, node ID j*(m+1)+i, two triangles per
// quad, sides ordered bottom, right, top, left.  Coordinates pass through "%.12f" text as they
// would through a written and re-read .14 file (:164).  Added here: seeded interior-vertex
// jitter and seeded Fourier-mode bathymetry for the large tidal workloads.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <memory>
#include <vector>

#include "../../include/rkdg.h"
#include "mesh_owner.hpp"


namespace {

// splitmix64: portable seeded stream (same numbers on every platform and compiler).
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

double through_text(double v) {  // "%.12f" then strtod, as a .14 round trip does
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.12f", v);
    return std::strtod(buf, nullptr);
}

struct Fourier {
    int n = 0;
    std::vector<double> kx, ky, ph, a;
    double mean = 0, lo = 0, hi = 0;
};

constexpr int kMaxModes = 4096;  // Fourier bathymetry modes

}  // namespace

extern "C" rkdg_status rkdg_mesh_generate(const rkdg_quad_mesh_spec* sp, rkdg_mesh** out, rkdg_mesh_desc* desc) {
    if (!sp || !out || !desc || sp->m < 1 || sp->n < 1 || sp->pattern < 0 || sp->pattern > 2) return RKDG_ERR_ARG;
    if (!(sp->x2 > sp->x1) || !(sp->y2 > sp->y1) || sp->bathy_kind < 0 || sp->bathy_kind > 3 ||
        sp->n_modes > kMaxModes)
        return RKDG_ERR_ARG;
    for (int s = 0; s < 4; ++s)
        if (sp->bc[s] < RKDG_BC_LAND || sp->bc[s] > RKDG_BC_FUNCTION) return RKDG_ERR_ARG;
    return rkdg::guarded("rkdg_mesh_generate", [&]() -> rkdg_status {
    std::unique_ptr<rkdg_mesh> owner(new rkdg_mesh());
    rkdg_mesh* M = owner.get();
    const int m = sp->m, n = sp->n;
    const double p1[2] = {sp->x1, sp->y1}, p2[2] = {sp->x2, sp->y1}, p3[2] = {sp->x2, sp->y2}, p4[2] = {sp->x1, sp->y2};
    const double dxi = 2.0 / m, deta = 2.0 / n;
    const size_t nn = (size_t)(m + 1) * (n + 1);
    M->xyz.resize(3 * nn);

    Rng rng(sp->seed);
    // Fourier bathymetry: depth = clamp(mean + sum_k a_k cos(kx_k x + ky_k y + ph_k), bmin, bmax),
    // wavelengths between 1/8 and 1 of the domain extent, amplitudes ~ 1/k.
    Fourier F;
    if (sp->bathy_kind == 2) {
        F.n = sp->n_modes > 0 ? sp->n_modes : 16;
        F.lo = sp->bmin, F.hi = sp->bmax, F.mean = 0.5 * (sp->bmin + sp->bmax);
        const double Lx = sp->x2 - sp->x1, Ly = sp->y2 - sp->y1, twopi = 6.283185307179586;
        double asum = 0;
        for (int k = 0; k < F.n; ++k) {
            const double cx = std::floor(rng.uniform() * 8.0) + 1.0, cy = std::floor(rng.uniform() * 8.0);
            const double sx = rng.uniform() < 0.5 ? -1.0 : 1.0;
            F.kx.push_back(sx * twopi * cx / Lx);
            F.ky.push_back(twopi * cy / Ly);
            F.ph.push_back(twopi * rng.uniform());
            const double a = (0.5 + rng.uniform()) / std::sqrt(cx * cx + cy * cy);
            F.a.push_back(a);
            asum += a;
        }
        const double scale = 0.75 * (F.hi - F.lo) / asum;  // keeps most of the field inside the clamp
        for (double& a : F.a) a *= scale;
    }
    auto depth = [&](double x, double y) -> double {
        switch (sp->bathy_kind) {
            case 0:
                return sp->b0;
            case 1:
                return -sp->alpha * (x * x + y * y);  // parabolic bowl (mesh_generators/bathymetry.hpp)
            case 2: {
                double b = F.mean;
                for (int k = 0; k < F.n; ++k) b += F.a[k] * std::cos(F.kx[k] * x + F.ky[k] * y + F.ph[k]);
                return b < F.lo ? F.lo : (b > F.hi ? F.hi : b);
            }
            case 3:
                return x < sp->step_x ? sp->step_b_left : sp->step_b_right;
            default:
                return sp->b0;
        }
    };
    const double hx = (sp->x2 - sp->x1) / m, hy = (sp->y2 - sp->y1) / n;
    for (int i = 0; i <= m; ++i)
        for (int j = 0; j <= n; ++j) {
            const double xi = -1.0 + dxi * i, eta = -1.0 + deta * j;
            const double N1 = (1.0 - xi) * (1.0 - eta) / 4.0, N2 = (1.0 + xi) * (1.0 - eta) / 4.0;
            const double N3 = (1.0 + xi) * (1.0 + eta) / 4.0, N4 = (1.0 - xi) * (1.0 + eta) / 4.0;
            double x = p1[0] * N1 + p2[0] * N2 + p3[0] * N3 + p4[0] * N4;
            double y = p1[1] * N1 + p2[1] * N2 + p3[1] * N3 + p4[1] * N4;
            const size_t id = (size_t)j * (m + 1) + i;
            if (sp->jitter > 0 && i > 0 && i < m && j > 0 && j < n) {
                // a generator seeded per node ID: the jitter of a node does not depend on loop order
                Rng r(sp->seed * 0x100000001b3ULL + id);
                x += (2 * r.uniform() - 1) * sp->jitter * hx;
                y += (2 * r.uniform() - 1) * sp->jitter * hy;
            }
            const double z = depth(x, y);
            M->xyz[3 * id] = through_text(x);
            M->xyz[3 * id + 1] = through_text(y);
            M->xyz[3 * id + 2] = through_text(z);
        }

    M->elt.resize((size_t)6 * m * n);
    auto set = [&](size_t e, int a, int b, int c) {
        M->elt[3 * e] = a, M->elt[3 * e + 1] = b, M->elt[3 * e + 2] = c;
    };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            const size_t e0 = (size_t)2 * i + (size_t)2 * m * j, e1 = e0 + 1;
            const int a = i + (m + 1) * j;
            bool flip;  // simple: every quad uses the default split; zigzag: odd rows; checker: alternate
            if (sp->pattern == 0)
                flip = false;
            else if (sp->pattern == 1)
                flip = (j % 2) == 1;
            else
                flip = ((i + j) % 2) == 1;
            if (!flip) {
                set(e0, a, a + m + 2, a + m + 1);
                set(e1, a, a + 1, a + m + 2);
            } else {
                set(e0, a, a + 1, a + m + 1);
                set(e1, a + 1, a + m + 2, a + m + 1);
            }
        }

    // boundary sides: bottom, right, top, left (quad_mesh_generator.cpp:150-163)
    std::vector<int> side[4];
    for (int i = 0; i <= m; ++i) side[0].push_back(i), side[2].push_back(i + (m + 1) * n);
    for (int i = 0; i <= n; ++i) side[1].push_back(m + i * (m + 1)), side[3].push_back(i * (m + 1));
    for (int s = 0; s < 4; ++s)
        for (size_t k = 0; k + 1 < side[s].size(); ++k) {
            M->bedge.push_back(side[s][k]);
            M->bedge.push_back(side[s][k + 1]);
            M->btype.push_back((uint8_t)sp->bc[s]);
        }
    bool forced = false;
    for (int s = 0; s < 4; ++s) forced |= (sp->bc[s] == RKDG_BC_TIDE || sp->bc[s] == RKDG_BC_FLOW);
    if (forced) {
        M->freq.push_back(sp->frequency);
        M->ff.push_back(sp->forcing_fact);
        M->eq.push_back(sp->equilib_arg);
        M->amp.assign(nn, 0.0);
        M->phase.assign(nn, 0.0);
        for (int s = 0; s < 4; ++s)
            if (sp->bc[s] == RKDG_BC_TIDE || sp->bc[s] == RKDG_BC_FLOW)
                for (int v : side[s]) M->amp[v] = sp->amplitude, M->phase[v] = sp->phase;
    }
    if (sp->manning_n > 0) M->manning.assign(nn, sp->manning_n);

    *desc = rkdg_mesh_desc{};
    M->fill(desc);
    *out = owner.release();
    return RKDG_OK;
    });
}

extern "C" void rkdg_mesh_free(rkdg_mesh* mesh) { delete mesh; }
