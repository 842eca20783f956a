// Khatri-Rao pair sketch kernels.
#pragma once

#include "common.hpp"

namespace mpisdf {

// Transposed Omega (formed once per u, v) and per-batch point scales.
struct SketchScratch {
    DeviceBuffer<double> ut, vt, scale;
    DeviceBuffer<double> part;  // per-row-group partial sums of ||y_g||^2
    const double* u_src = nullptr;
    const double* v_src = nullptr;
    int64_t ell = 0, d = 0;
    LaunchTimer timer;  // sketch kernel launches
};

// Y[g * ldy + s] for g < count, s < ell; phi is [count][d] FP64, u and v are
// [d][ell] row-major (u[a * ell + s]); sw[g] = sqrt(|w_g|).  Non-finite
// outputs are counted into *bad; ynorm2[g] += ||y_g||^2 when given.
void sketch_f64(const double* phi, int64_t count, int64_t d, const double* u, const double* v, int64_t ell,
                const double* sw, double* y, int64_t ldy, SketchScratch& scratch, unsigned long long* bad,
                hipStream_t stream, double* ynorm2 = nullptr);

}  // namespace mpisdf
