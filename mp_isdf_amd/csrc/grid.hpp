// Molecular integration grid without PySCF (see grid.cpp).
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace mpisdf {

// Returns (xyz[3 n], weights[n]) in Bohr for atoms with nuclear charges
// ``charges`` at ``coords`` (Bohr); ``level`` follows PySCF's 0..9 scale.
// ``device``: partition on the GPU when one is present (identical weights).
std::pair<std::vector<double>, std::vector<double>> becke_grid(const int* charges, const double* coords,
                                                               int64_t natm, int level, bool device = true);
// Number of points becke_grid returns, without computing them.
int64_t becke_grid_size(const int* charges, int64_t natm, int level);

}  // namespace mpisdf
