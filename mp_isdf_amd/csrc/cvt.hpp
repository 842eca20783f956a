// Centroidal Voronoi anchor selection (see cvt.hip).
#pragma once

#include "common.hpp"

#include <vector>

namespace mpisdf {

// Sorted distinct original indices of `rank` sites chosen by mass-weighted
// k-means++ seeding, Lloyd iterations and snapping (cpp/src/anchor_select_cvt.cpp:17-204).
// Host arrays points[npts][3], weights[npts], importance[npts]; the mass of a
// point is its importance where its quadrature weight is nonzero.
std::vector<int64_t> select_cvt(const double* points, const double* weights, const double* importance, int64_t npts,
                                int64_t rank, uint64_t seed, int64_t max_iter, double tol, hipStream_t s);

}  // namespace mpisdf
