#pragma once

#include "common.hpp"

namespace mpisdf {

// Dense weighted pair matrix, m = w^2 rows per grid column (column-major).
void launch_pair_product(const double* f, int64_t n, int64_t w, const double* sw, double* a,
                         unsigned long long* bad, hipStream_t s);

// out[j] += sum_g f[g * w + j]^2 (out must be initialised); the rows are
// summed into per-slot partials (part, resized as needed) that are added in
// slot order, so the result is the same on every run.
void launch_colsumsq(const double* f, int64_t n, int64_t w, double* out, DeviceBuffer<double>& part, hipStream_t s);

// Fixed-order reductions.  Every FP64 sum over workgroups in this library is
// written as per-workgroup partials and added by one of these in index order,
// never by atomics in completion order, so a result is bit-identical from run
// to run.
// out[j] = (accumulate ? out[j] : 0) + sum_{r < groups} part[r * len + j], r ascending.
void launch_sum_partials(const double* part, int64_t groups, int64_t len, double* out, bool accumulate,
                         hipStream_t s);
// out[c] += sum_{i < n} x[c * ld + i] for c < nvec, each in a fixed order.
void launch_sum_fixed(const double* x, int64_t n, int64_t ld, int64_t nvec, double* out, hipStream_t s);

// out[j] = ||y_j||^2 (FP64) of the FP32 columns y[j * ldy + i], i < m.
void launch_column_norms2_f32(const float* y, int64_t m, int64_t n, int64_t ldy, double* out, hipStream_t s);
// The same for FP64 columns.
void launch_column_norms2_f64(const double* y, int64_t m, int64_t n, int64_t ldy, double* out, hipStream_t s);

}  // namespace mpisdf
