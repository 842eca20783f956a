// Device kernels of the online QRCP (see online_kernels.hip).
#pragma once

#include "common.hpp"

namespace mpisdf {

// norms[c] = || (H[rank-1]..H[0] a_{ids[c]})[first:] ||; out (optional) gets
// the transformed columns.  scratch holds scratch_columns * d^2 doubles.
void online_transform_norms(const double* feat, const double* sw, int64_t d, const int64_t* ids, int64_t count,
                            const double* refl, const double* tau, int64_t rank, int64_t first, double* scratch,
                            int64_t scratch_columns, double* norms, double* out, hipStream_t s);

// q = H[0] ... H[k] e_k (m = d^2).
void online_make_q(const double* refl, const double* tau, int64_t m, int64_t k, double* q, hipStream_t s);

void online_abs(const double* x, int64_t count, double* y, hipStream_t s);

// row[g] = q^T a_g, absrow[g] = sum_i |q_i a_gi| for all n columns (quadratic forms).
void online_row_pass(const double* feat, const double* absfeat, const double* sw, int64_t n, int64_t d,
                     const double* q, double* absq, double* t, double* row, double* absrow, hipStream_t s);

}  // namespace mpisdf
