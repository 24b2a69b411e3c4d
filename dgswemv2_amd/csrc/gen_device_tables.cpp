// Build-time generator of device_tables.h: the reference-element tables of p = 1, 2, 3 as
// compile-time constants (exact hex-float literals of the host FP64 tables from tables.cpp),
// so the kernels read them as instruction constants instead of memory.
#include <cstdio>
#include <string>
#include <vector>

#include "tables.hpp"

static void emit(const char* name, const std::vector<double>& v, int rows, int cols, const char* type) {
    std::printf("    static constexpr %s %s[%d][%d] = {\n", type, name, rows, cols);
    for (int r = 0; r < rows; ++r) {
        std::printf("        {");
        for (int c = 0; c < cols; ++c) {
            const double x = v[(size_t)r * cols + c];
            if (std::string(type) == "float")
                std::printf("%a%s", (double)(float)x, c + 1 < cols ? ", " : "");
            else
                std::printf("%a%s", x, c + 1 < cols ? ", " : "");
        }
        std::printf("},\n");
    }
    std::printf("    };\n");
}

static void emit_type(int p, const char* type) {
    rkdg::Master m = rkdg::make_master(p);
    const int nd = m.ndof, ng = m.ngp, ne = m.ngp_e;
    std::printf("template <>\nstruct ConstTables<%s, %d> {\n", type, p);
    emit("phi", m.phi, nd, ng, type);
    emit("dphi0", m.dphi[0], nd, ng, type);
    emit("dphi1", m.dphi[1], nd, ng, type);
    emit("idphi0", m.idphi[0], ng, nd, type);
    emit("idphi1", m.idphi[1], ng, nd, type);
    emit("ipf", m.ipf, ng, nd, type);
    emit("minv", m.minv, 1, nd, type);
    emit("we", m.erule.w, 1, ne, type);
    std::vector<double> phie, psi(3 * (size_t)ng), psie;
    for (int f = 0; f < 3; ++f) phie.insert(phie.end(), m.phi_e[f].begin(), m.phi_e[f].end());
    emit("phie", phie, 3 * nd, ne, type);  // row f*nd + d
    for (int g = 0; g < ng; ++g) {
        const double z1 = m.rule.z[g][0], z2 = m.rule.z[g][1];
        psi[g] = -(z1 + z2) / 2, psi[ng + g] = (1 + z1) / 2, psi[2 * ng + g] = (1 + z2) / 2;
    }
    emit("psi", psi, 3, ng, type);
    for (int f = 0; f < 3; ++f)
        for (int node = 0; node < 3; ++node)
            for (int g = 0; g < ne; ++g) {
                const double z1 = m.zedge[f][g][0], z2 = m.zedge[f][g][1];
                psie.push_back(node == 0 ? -(z1 + z2) / 2 : (node == 1 ? (1 + z1) / 2 : (1 + z2) / 2));
            }
    emit("psie", psie, 9, ne, type);  // row f*3 + node
    // edge weights folded into the edge basis: phie(d, g) * w(g) (FP32 fast form)
    std::vector<double> phiew;
    for (int f = 0; f < 3; ++f)
        for (int d = 0; d < nd; ++d)
            for (int g = 0; g < ne; ++g) phiew.push_back(m.phi_e[f][(size_t)d * ne + g] * m.erule.w[g]);
    emit("phiew", phiew, 3 * nd, ne, type);
    std::printf("};\n\n");
}

int main() {
    std::printf("// Generated at build time by gen_device_tables (from tables.cpp) -- do not edit.\n");
    std::printf("#pragma once\nnamespace rkdg {\ntemplate <typename Real, int P>\nstruct ConstTables;\n\n");
    for (int p = 1; p <= 3; ++p) {
        emit_type(p, "double");
        emit_type(p, "float");
    }
    std::printf("}  // namespace rkdg\n");
    return 0;
}
