// Dense audit of sketch pivots (cpp/src/anchor_select_randomized_id.cpp:19-88).
#pragma once

#include <vector>

#include "common.hpp"

namespace mpisdf {

// Relative Frobenius residual ||A - Q Q^T A||_F / ||A||_F of the m x n
// column-major device matrix A (leading dimension m) after projecting every
// column on the span of the columns listed in pivots.  Q is built from those
// columns in order by two-pass Gram-Schmidt; a column whose remaining norm is
// at or below 32 eps of its own norm adds no direction, as in the reference.
// The projection is applied twice.  Returns 0 for a zero matrix.
double projection_relative_error(const double* a, int64_t m, int64_t n, const std::vector<int64_t>& pivots,
                                 hipStream_t s);

}  // namespace mpisdf
