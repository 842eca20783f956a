// Robust blockwise random pivoting on AO pair rows (see rbrp.cpp).
#pragma once

#include "common.hpp"

#include <limits>
#include <string>
#include <vector>

namespace mpisdf {

struct RbrpOptions {
    int64_t batch_size = 128;  // accepted for the reference's signature; the GPU forms all rows at once
    int64_t block_size = 16;
    double tolerance = 1e-10;
    int64_t max_rank = 0;
    uint64_t seed = 7;
    double tau_block = std::numeric_limits<double>::quiet_NaN();  // NaN: 1 / block_size
    double rank_rtol = 1e-12;
    int64_t refresh_every = 8;
    double svd_rtol = 1e-12;
    bool compute_id = false;
};

struct RbrpResult {
    std::vector<int64_t> indices;      // rows of the AO matrix, in selection order
    std::vector<double> interp_vecs;   // [rank][n] when compute_id
    std::vector<double> transform;     // [rank][nao * nao] when compute_id
    std::vector<double> history;       // [blocks][4]: rank, candidates, retained, relative residual
    double relative_residual = 0, condition = 0;
    int64_t retained_rank = 0;
    std::string status;
};

// ao_host: [n][nao] row-major on the host (always needed: the block rows and
// validation read it); ao_dev: the same on the device, or null to upload it.
RbrpResult rbrp_ao(const double* ao_host, const double* ao_dev, const double* weights, int64_t n, int64_t nao,
                   const RbrpOptions& opt, hipStream_t s);

}  // namespace mpisdf
