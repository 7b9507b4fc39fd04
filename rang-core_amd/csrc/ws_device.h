// The w-stacking gridder on the GPU (core/src/grid.rs), behind the C ABI of
// include/rang_core_grid.h.
#pragma once

#include <cstddef>
#include <string>

#include "../../include/rang_core_grid.h"

namespace rang {
namespace ws {

// Inversion on the current HIP device; grid and transforms in single or
// double precision. 0 ok, 1 invalid arguments, 2 device failure (message in
// err). timings: 14 doubles or null.
int invert_gpu(const GridInputs& in, const GridOutputs& out, bool single, double* timings,
               std::string& err);

// The oversampled Kaiser-Bessel kernel (kernel_size * oversampling values).
int kernel(std::size_t kernel_size, std::size_t oversampling, double* out);

}  // namespace ws
}  // namespace rang
