// Khatri-Rao pair sketch kernels.
#pragma once

#include "common.hpp"

namespace mpisdf {

struct SketchScratch {
    DeviceBuffer<double> left, right;
};

// Y[g * ldy + s] for g < count, s < ell; phi is [count][d] FP64, u and v are
// [d][ell] row-major (u[a * ell + s]); sw[g] = sqrt(|w_g|).  Non-finite
// outputs are counted into *bad.
void sketch_f64(const double* phi, int64_t count, int64_t d, const double* u, const double* v, int64_t ell,
                const double* sw, double* y, int64_t ldy, SketchScratch& scratch, unsigned long long* bad,
                hipStream_t stream);

}  // namespace mpisdf
