// Integral-direct pivoted Cholesky of the AO-pair Coulomb matrix and the
// Coulomb metric through its factors (see eri.hip).
#pragma once

#include "context.hpp"

#include <vector>

namespace mpisdf {

enum class CholeskyAlgorithm { UnblockedGreedy = 0, UnblockedUniform = 1 };

// Packed AO pairs x >= y, index x (x + 1) / 2 + y.  factors is row-major
// [rank][npair] on the device; pivots are packed pair indices in selection
// order; residual[npair] is the final residual diagonal (device).
struct EriCholesky {
    int64_t npair = 0, rank = 0;
    DeviceBuffer<double> factors;
    DeviceBuffer<double> residual;
    std::vector<int64_t> pivots;
};

// cpp/src/cholesky_eri_greedy.cpp:53-175 (both algorithms), with
// tolerance relative to the largest initial diagonal and max_rank 0 for no
// cap.  s, p and d shells.
void eri_cholesky(Context& ctx, double tolerance, int64_t max_rank, CholeskyAlgorithm algorithm, uint64_t seed,
                  bool debug, EriCholesky& out, hipStream_t s);

// out[nvec][nvec] (host) = P P^T with P = packed(transform) F^T, transform
// row-major [nvec][nao * nao] on the host (cpp/src/coulomb_metric.cpp:20-98).
void eri_coulomb_metric(Context& ctx, const double* transform, int64_t nvec, CholeskyAlgorithm algorithm,
                        double tolerance, int64_t max_rank, int64_t batch_size, bool debug, uint64_t seed,
                        double* out, hipStream_t s);

}  // namespace mpisdf
