// Column-pivoted QR (greedy Businger-Golub pivot order) on the GPU.
#pragma once

#include "common.hpp"

namespace mpisdf {

struct QrcpOptions {
    int64_t kmax = 0;            // maximum pivots (<= min(m, n))
    double stop_abs = 0.0;       // stop before a pivot whose norm is <= stop_abs
    double stop_rel = 0.0;       // ... or <= stop_rel * |R_00|
    int64_t candidates = 1024;   // explicit FP64 residual columns per block
    double margin = 0.0;         // relative error bound of the folded norms
    bool want_energy = false;    // row energies sum_j (q_i^T y_j)^2 (randomized tail rule)
    bool want_tail = false;      // exact residual Frobenius^2 after the last pivot
    bool debug = false;
    const double* ynorm2 = nullptr;  // optional device ||y_j||^2 (skips the norm pass; required for kI8Digits)
    const double* yscale = nullptr;  // kI8Digits: device per-column scales of the digit planes
    bool split_bf16_fold = false;    // FP32 Y: fold with split-BF16 MFMA (needs margin > 0)
    // Leave columns far below the pivots unfolded: a column's key is then its
    // ||y_j||^2, still an upper bound of its residual.  Only the prefix whose
    // largest ||y_j||^2 can reach 1% of the last certification bound is
    // folded; once the unfolded suffix passes 10% of it the prefix widens, and
    // the new columns get one fold against every pivot so far.  Pivots are
    // unchanged; the columns should come in roughly descending ||y_j||^2.
    // Not with energy or tail.
    bool dormant = false;
    // Optional device ranks (a permutation of 0 .. n-1): ties between equal
    // keys or norms go to the lowest rank instead of the lowest column.
    const int64_t* tie_rank = nullptr;
};

struct QrcpResult {
    std::vector<int64_t> pivots;  // 0-based columns in pivot order
    std::vector<double> rdiag;    // |R_ii|
    std::vector<double> energy;   // per pivot row (if requested)
    double frob2 = 0.0;           // ||Y||_F^2
    double tail2 = 0.0;           // residual Frobenius^2 after the last pivot (if requested)
    double fold_seconds = 0.0;    // summed HIP-event duration of the fold launches
    double fold_flops = 0.0;      // their algorithmic flops, sum 2 b m n
    double gs_seconds = 0.0;      // summed HIP-event duration of the candidate Gram-Schmidt products
    double gs_flops = 0.0;        // their flops: 4 k m cols per pass (H = Q^T R and R -= Q H)
    int64_t folded_columns = 0;   // columns folded at the end (dormant: the active prefix)
    int64_t blocks = 0, folds = 0, stalls = 0, refreshes = 0;
};

// Y is m x n column-major with leading dimension ldy, element type f64 or f32;
// or, for a digit type, column j at byte offset digit_column_bytes(ytype) ldy j
// holding digit_planes_of(ytype) signed digits per entry (digit_offset;
// y_hat_j = yscale[j] sum_r d_r B^-(r+1), B = digit_base_of(ytype)), folded
// with INT8 MFMA and factorised as y_hat (candidate columns are rebuilt from
// the same digits; opt.margin must cover the fold's error).  Y is read only.
// Pivot decisions use exact FP64 residual norms of the candidate columns
// (LAPACK dgeqp3 pivot semantics: max norm, lowest index on ties, or lowest
// tie rank when opt.tie_rank is given).
QrcpResult qrcp_device(const void* y, DType ytype, int64_t m, int64_t n, int64_t ldy,
                       const QrcpOptions& opt, hipStream_t stream);

}  // namespace mpisdf
