// Output writers (writers.cpp): the reference's modal text dump, VTU and VTK legacy files.
#pragma once

#include <cstdint>
#include <string>

#include "model.hpp"

namespace rkdg {

// q: host state [3][ndof][ne] (FP64); t: stage time; step: the stepper's step counter.
bool write_modal(const Model& M, const double* q, double t, int step, const std::string& prefix, std::string& err);
bool write_vtu(const Model& M, const double* q, const uint8_t* wet, int step, const std::string& prefix,
               std::string& err);
bool write_vtk(const Model& M, const double* q, const uint8_t* wet, int step, const std::string& prefix,
               std::string& err);

}  // namespace rkdg
