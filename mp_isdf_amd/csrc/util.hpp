#pragma once

#include "common.hpp"

namespace mpisdf {

// Dense weighted pair matrix, m = w^2 rows per grid column (column-major).
void launch_pair_product(const double* f, int64_t n, int64_t w, const double* sw, double* a,
                         unsigned long long* bad, hipStream_t s);

// out[j] += sum_g f[g * w + j]^2 (out must be initialised).
void launch_colsumsq(const double* f, int64_t n, int64_t w, double* out, hipStream_t s);

// out[j] = ||y_j||^2 (FP64) of the FP32 columns y[j * ldy + i], i < m.
void launch_column_norms2_f32(const float* y, int64_t m, int64_t n, int64_t ldy, double* out, hipStream_t s);

}  // namespace mpisdf
