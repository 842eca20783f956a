// Anchor selection drivers (device counterparts of the reference's
// AnchorSelection::select and its three local algorithms).
#pragma once

#include "axis.hpp"
#include "context.hpp"

namespace mpisdf {

// LocalAnchorSelectionAlgorithm values (cpp/include/enums.hpp:23-30); 3 is CVT.
enum Algorithm : int { kDense = 0, kOnline = 1, kRandomized = 2, kDensity = 4, kDensityNormalized = 5 };
// kPrecF64I8: the FP64 sketch, factorised as four base-254 INT8 digit planes
// (the INT8 fold in place of the FP64 one; the same pivots as kPrecF64 unless
// two candidates tie to ~1e-10).
enum Precision : int { kPrecF64 = 0, kPrecF32 = 1, kPrecBF16 = 2, kPrecFP8 = 3, kPrecBF16x3 = 4, kPrecF64I8 = 5 };

struct SelectArgs {
    const double* points = nullptr;   // host [n][3]
    const double* weights = nullptr;  // host [n]
    int64_t n = 0;
    std::vector<int32_t> axes;
    const double* mo_coeff = nullptr;
    int64_t nocc = 0;
    std::vector<double> axis_weights;
    double tolerance = 1e-10;
    int64_t max_rank = 0;
    int algorithm = kDense;
    int64_t batch_size = 128;
    bool debug = false;
    int64_t sketch_size = 128;
    int64_t oversampling = 16;
    uint64_t seed = 7;
    int precision = kPrecF64;
    // Density Khatri-Rao sketches (kDensity, kDensityNormalized): host [n]
    // nonnegative importance, and the AO squared-norm screening threshold.
    const double* importance = nullptr;
    double ao_screening_threshold = 1e-24;
    // rand_sketch and the density sketches: after selection, compare the
    // pivots' projection residual of the actual weighted pair matrix with
    // that of equally many dense QRCP pivots
    // (cpp/src/anchor_select_randomized_id.cpp:19-88, :164-165, :389-400).
    bool audit = false;
};

// Diagnostics slots (double[16]) written when requested.
enum Diag : int {
    kDiagRows = 0,          // sketch rows ell (randomized) or pair rows d^2
    kDiagColumns = 1,       // kept grid columns
    kDiagNumericalRank = 2,
    kDiagCap = 3,
    kDiagRank = 4,
    kDiagNorm = 5,          // sketch Frobenius norm (randomized)
    kDiagRelTail = 6,       // tails[rank] / norm (tolerance mode)
    kDiagBlocks = 7,
    kDiagFolds = 8,
    kDiagSecondsSketch = 9,
    kDiagSecondsQrcp = 10,
    kDiagSecondsTotal = 11,
    kDiagStalls = 12,       // QRCP blocks without a certified pivot
    kDiagRefreshes = 13,    // exact FP64 refreshes of the folded norms
    kDiagFoldSeconds = 14,  // summed duration of the QRCP fold launches (HIP events)
    kDiagFoldFlops = 15,    // their algorithmic FP64 flops: sum 2 b ell n
    kDiagFoldLaunches = 16,
    kDiagSketchSeconds = 17,  // summed duration of the sketch kernel launches (HIP events)
    kDiagSketchFlops = 18,    // their MFMA flops: sum 4 ell d count per product term (3 split BF16, 2 FP8)
    kDiagSketchLaunches = 19,
    kDiagFoldKind = 20,       // 0 FP64 MFMA on FP64 Y, 1 FP64 MFMA on FP32 Y, 2 / 3 INT8 MFMA on 3 / 4 base-128 planes, 4 on 4 base-254 planes
    kDiagSketchOps = 21,      // fp64i8: INT8 operations of the digit sketch (15 pairs x 2 (2 ell64) d64 per point); else 0
    kDiagGsSeconds = 22,      // summed duration of the QRCP's candidate Gram-Schmidt products (HIP events)
    kDiagGsFlops = 23,        // their FP64 flops: 4 k m cols per pass
    // Audit (SelectArgs::audit): relative projection residuals of the selected
    // and of the dense QRCP pivots, quadrature columns sqrt|w| pair_g, then,
    // for the density sketches, the sketch's own columns; -1 when not run.
    kDiagAuditActual = 24,
    kDiagAuditDense = 25,
    kDiagAuditDensityActual = 26,
    kDiagAuditDensityDense = 27,
    kDiagCount = 28
};

std::vector<int64_t> select_anchors(Context& ctx, const SelectArgs& args, double* diagnostics);

// Gaussian factors (cpp/src/anchor_select_randomized_id.cpp:269-279), host.
void make_omega(int64_t d, int64_t ell, uint64_t seed, std::vector<double>& u, std::vector<double>& v);

}  // namespace mpisdf
