// Output writers: the reference's modal text dump and VTU file, from the host copy of a state.
//
// Modal (write_modal_data, problem_postprocessor/swe_post_write_modal.hpp:5-90): four files
// <prefix>_modal_{ze,qx,qy,bath}.txt, truncated at step 0 and appended to afterwards; each
// record is the stage time ("%f") then one line "ID value" ("%.6e") per element (ascending ID)
// and mode.
//
// VTU (write_VTU_data, swe_post_write_vtu.hpp:5-103, with the geometry head and foot of
// Writer::InitializeMeshGeometryVTU, simulation/writer.hpp:260-350, merged as
// Writer::WriteOutput does, :153-190): <prefix>_data_<step>.vtu, an ASCII UnstructuredGrid with
// one linear cell per element (the reference's N_DIV = 1 post-processing subdivision,
// general_definitions.hpp:229): its three vertices as points (master coordinates (-1,-1),
// (1,-1), (-1,1), StraightTriangle::GetVTK, shape_straighttriangle.cpp:296-350), point data
// ze/qx/qy/bath and |q|/h at the vertices, cell data ze/qx/qy/bath at the centroid (-1/3,-1/3)
// (Triangle::VTKPostCell / VTKPostPoint, master_triangle.tpp:155-195), the element ID and the
// wet flag.  went_completely_dry is set by the reference's wetting-drying pass when the mean
// vertex depth is at or below h_o's threshold (swe_post_wet_dry.hpp:24-28), which leaves the
// element at exactly that mean depth; it is written here as that condition on the state.
//
// VTK legacy (write_VTK_data, swe_post_write_vtk.hpp:5-95, after the geometry of
// Writer::InitializeMeshGeometryVTK, simulation/writer.hpp:195-258, merged as WriteOutput does,
// :121-151): <prefix>_data_<step>.vtk, ASCII, the same points, cells and values as the VTU (cell
// data ze/qx/qy/bath, ID, wet_dry, went_completely_dry; point data ze/qx/qy/bath), in the
// reference's layout and its iostream number formatting (six significant digits, values rounded
// to float first).
#include <array>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "model.hpp"
#include "tables.hpp"
#include "writers.hpp"

namespace rkdg {

namespace {

struct File {
    FILE* f = nullptr;
    explicit File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// value of mode-expanded data u[dof] (stride `stride`, element e) at basis values phi[dof]
double at(const double* u, size_t stride, int e, int nd, const double* phi) {
    double s = 0.0;
    for (int d = 0; d < nd; ++d) s += u[(size_t)d * stride + e] * phi[d];
    return s;
}

// Post-processing values of one state: the modal expansions evaluated at the three vertices and
// the centroid of each element (the reference's N_DIV = 1 subdivision).
struct PostValues {
    const Model& M;
    const double* q;
    int nd;
    size_t ne;
    std::vector<double> phi;  // [nd][4]: vertices 0-2, centroid 3
    std::vector<double> pv;
    PostValues(const Model& M_, const double* q_) : M(M_), q(q_), nd(M_.ndof), ne((size_t)M_.ne) {
        const std::vector<std::array<double, 2>> pts = {{-1.0, -1.0}, {1.0, -1.0}, {-1.0, 1.0},
                                                         {-1.0 / 3.0, -1.0 / 3.0}};
        phi.resize((size_t)nd * 4);
        dubiner_eval(M.p, pts, phi.data(), nullptr, nullptr);
        pv.resize(nd);
    }
    const double* phi_at(int pt) {
        for (int d = 0; d < nd; ++d) pv[d] = phi[(size_t)d * 4 + pt];
        return pv.data();
    }
    double var(int v, int e, int pt) { return at(q + (size_t)v * nd * ne, ne, e, nd, phi_at(pt)); }
    double bath(int e, int pt) { return at(M.bath.data(), ne, e, nd, phi_at(pt)); }
    // mean vertex depth of the linear part (modes 0-2 at the vertices) + vertex bathymetry at or
    // below the wetting-drying threshold
    bool went_completely_dry(int e) {
        double hsum = 0.0;
        for (int pt = 0; pt < 3; ++pt) hsum += at(q, ne, e, nd < 3 ? nd : 3, phi_at(pt)) + M.bath_vrtx[(size_t)pt * ne + e];
        return hsum / 3.0 <= M.h_thr;
    }
};

}  // namespace

bool write_vtk(const Model& M, const double* q, const uint8_t* wet, int step, const std::string& prefix,
               std::string& err) {
    const int ne = M.ne;
    const std::string path = prefix + "_data_" + std::to_string(step) + ".vtk";
    File fl(path, "w");
    if (!fl.f) {
        err = "cannot open " + path;
        return false;
    }
    FILE* f = fl.f;
    PostValues P(M, q);
    std::fprintf(f, "# vtk DataFile Version 3.0\nOUTPUT DATA\nASCII\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(f, "POINTS %d double\n", 3 * ne);
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k) std::fprintf(f, "%g\t%g\t%g\n", M.vxy[(size_t)k * ne + e], M.vxy[(size_t)(3 + k) * ne + e], 0.0);
    std::fprintf(f, "CELLS %d %d\n", ne, 4 * ne);
    for (int e = 0; e < ne; ++e) std::fprintf(f, "3\t%d\t%d\t%d\t\n", 3 * e, 3 * e + 1, 3 * e + 2);
    std::fprintf(f, "CELL_TYPES %d\n", ne);
    for (int e = 0; e < ne; ++e) std::fprintf(f, "5\n");  // straight_triangle (general_definitions.hpp:242)
    std::fprintf(f, "CELL_DATA %d\n", ne);
    const char* vn[3] = {"ze", "qx", "qy"};
    for (int v = 0; v < 3; ++v) {
        std::fprintf(f, "SCALARS %s_cell double 1\nLOOKUP_TABLE default\n", vn[v]);
        for (int e = 0; e < ne; ++e) std::fprintf(f, "%g\n", (double)(float)P.var(v, e, 3));
    }
    std::fprintf(f, "SCALARS bath_cell double 1\nLOOKUP_TABLE default\n");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%g\n", (double)(float)P.bath(e, 3));
    std::fprintf(f, "SCALARS ID unsigned_int 1\nLOOKUP_TABLE default\n");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d\n", e);
    std::fprintf(f, "SCALARS wet_dry bit 1\nLOOKUP_TABLE default\n");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d\n", wet[e] ? 1 : 0);
    std::fprintf(f, "SCALARS went_completely_dry bit 1\nLOOKUP_TABLE default\n");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d\n", P.went_completely_dry(e) ? 1 : 0);
    std::fprintf(f, "POINT_DATA %d\n", 3 * ne);
    for (int v = 0; v < 3; ++v) {
        std::fprintf(f, "SCALARS %s_point double 1\nLOOKUP_TABLE default\n", vn[v]);
        for (int e = 0; e < ne; ++e)
            for (int pt = 0; pt < 3; ++pt) std::fprintf(f, "%g\n", (double)(float)P.var(v, e, pt));
    }
    std::fprintf(f, "SCALARS bath_point double 1\nLOOKUP_TABLE default\n");
    for (int e = 0; e < ne; ++e)
        for (int pt = 0; pt < 3; ++pt) std::fprintf(f, "%g\n", (double)(float)P.bath(e, pt));
    if (std::ferror(f)) {
        err = "write error on " + path;
        return false;
    }
    return true;
}

bool write_modal(const Model& M, const double* q, double t, int step, const std::string& prefix, std::string& err) {
    const int ne = M.ne, nd = M.ndof;
    const char* names[4] = {"ze", "qx", "qy", "bath"};
    for (int v = 0; v < 4; ++v) {
        const std::string path = prefix + "_modal_" + names[v] + ".txt";
        File fl(path, step == 0 ? "w" : "a");
        if (!fl.f) {
            err = "cannot open " + path;
            return false;
        }
        std::fprintf(fl.f, "%f\n", t);
        for (int e = 0; e < ne; ++e)
            for (int d = 0; d < nd; ++d) {
                const double x = v < 3 ? q[((size_t)v * nd + d) * ne + e] : M.bath[(size_t)d * ne + e];
                std::fprintf(fl.f, "%d %.6e\n", e, x);
            }
        if (std::ferror(fl.f)) {
            err = "write error on " + path;
            return false;
        }
    }
    return true;
}

bool write_vtu(const Model& M, const double* q, const uint8_t* wet, int step, const std::string& prefix,
               std::string& err) {
    const int ne = M.ne;
    const std::string path = prefix + "_data_" + std::to_string(step) + ".vtu";
    File fl(path, "w");
    if (!fl.f) {
        err = "cannot open " + path;
        return false;
    }
    FILE* f = fl.f;
    PostValues P(M, q);
    auto qv = [&](int var, int e, int pt) { return P.var(var, e, pt); };
    auto bv = [&](int e, int pt) { return P.bath(e, pt); };

    std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\">\n\t<UnstructuredGrid>\n");
    std::fprintf(f, "\t\t<Piece NumberOfPoints=\"%d\" NumberOfCells=\"%d\">\n", 3 * ne, ne);
    std::fprintf(f, "\t\t\t<PointData>\n");
    const char* vn[3] = {"ze", "qx", "qy"};
    for (int v = 0; v < 3; ++v) {
        std::fprintf(f, "\t\t\t\t<DataArray type=\"Float32\" Name=\"%s_point\" format=\"ascii\">\n\t\t\t\t\t", vn[v]);
        for (int e = 0; e < ne; ++e)
            for (int pt = 0; pt < 3; ++pt) std::fprintf(f, "%g ", (double)(float)qv(v, e, pt));
        std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    }
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Float32\" Name=\"bath_point\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e)
        for (int pt = 0; pt < 3; ++pt) std::fprintf(f, "%g ", (double)(float)bv(e, pt));
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Float32\" Name=\"velocity_point\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e)
        for (int pt = 0; pt < 3; ++pt)
            std::fprintf(f, "%g ", (double)(float)(std::hypot(qv(1, e, pt), qv(2, e, pt)) / (qv(0, e, pt) + bv(e, pt))));
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n\t\t\t</PointData>\n\t\t\t<CellData>\n");
    for (int v = 0; v < 3; ++v) {
        std::fprintf(f, "\t\t\t\t<DataArray type=\"Float32\" Name=\"%s_cell\" format=\"ascii\">\n\t\t\t\t\t", vn[v]);
        for (int e = 0; e < ne; ++e) std::fprintf(f, "%g ", (double)(float)qv(v, e, 3));
        std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    }
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Float32\" Name=\"bath_cell\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%g ", (double)(float)bv(e, 3));
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"UInt32\" Name=\"ID\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d ", e);
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Int8\" Name=\"wet_dry\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d ", wet[e] ? 1 : 0);
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Int8\" Name=\"went_completely_dry\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d ", P.went_completely_dry(e) ? 1 : 0);
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n\t\t\t</CellData>\n");
    // geometry (the reference's "foot")
    std::fprintf(f, "\t\t\t<Points>\n\t\t\t\t<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e)
        for (int k = 0; k < 3; ++k)
            std::fprintf(f, "%g %g %g ", M.vxy[(size_t)k * ne + e], M.vxy[(size_t)(3 + k) * ne + e], 0.0);
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n\t\t\t</Points>\n\t\t\t<Cells>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d %d %d ", 3 * e, 3 * e + 1, 3 * e + 2);
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d ", 3 * (e + 1));
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n");
    std::fprintf(f, "\t\t\t\t<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n\t\t\t\t\t");
    for (int e = 0; e < ne; ++e) std::fprintf(f, "%d ", 5);  // VTK_TRIANGLE
    std::fprintf(f, "\n\t\t\t\t</DataArray>\n\t\t\t</Cells>\n\t\t</Piece>\n\t</UnstructuredGrid>\n</VTKFile>\n");
    if (std::ferror(f)) {
        err = "write error on " + path;
        return false;
    }
    return true;
}

}  // namespace rkdg
