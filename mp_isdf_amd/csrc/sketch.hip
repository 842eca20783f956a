// Khatri-Rao pair sketch Y[s, g] = ell^{-1/2} sqrt(h_g) (u_s . z_g)(v_s . z_g)
// (cpp/src/anchor_select_randomized_id.cpp:281-309), column-major Y[g*ldy + s].
//
// FP64 form: U^T and V^T (ell x d, K-contiguous) are formed once per Omega;
// each batch is one TN MFMA launch whose epilogue multiplies the two
// projections, scales by sqrt(h_g / ell), writes Y, sums ||y_g||^2 per row
// group (added in group order afterwards) and counts non-finite entries (the reference's id_check_sketch, :310-317).
#include "kernels.hpp"
#include "sketch.hpp"

#include <algorithm>
#include <cmath>

namespace mpisdf {
namespace {

// out[c * rows + r] = in[r * cols + c]
__global__ void transpose_f64_kernel(const double* __restrict__ in, int64_t rows, int64_t cols,
                                     double* __restrict__ out)
{
    __shared__ double tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int64_t rr = r0 + r, cc = c0 + tx;
        tile[r][tx] = (rr < rows && cc < cols) ? in[rr * cols + cc] : 0.0;
    }
    __syncthreads();
    for (int c = ty; c < 32; c += 8) {
        const int64_t cc = c0 + c, rr = r0 + tx;
        if (rr < rows && cc < cols) out[cc * rows + rr] = tile[tx][c];
    }
}

__global__ void point_scale_kernel(const double* __restrict__ sw, int64_t count, double inv_sqrt_ell,
                                   double* __restrict__ scale)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < count) scale[g] = sw[g] * inv_sqrt_ell;
}

}  // namespace

void sketch_f64(const double* phi, int64_t count, int64_t d, const double* u, const double* v, int64_t ell,
                const double* sw, double* y, int64_t ldy, SketchScratch& scratch, unsigned long long* bad,
                hipStream_t stream, double* ynorm2)
{
    if (count <= 0) return;
    if (scratch.u_src != u || scratch.v_src != v || scratch.ell != ell || scratch.d != d) {
        scratch.ut.resize((size_t)(ell * d));
        scratch.vt.resize((size_t)(ell * d));
        const dim3 grid((unsigned)((ell + 31) / 32), (unsigned)((d + 31) / 32));
        hipLaunchKernelGGL(transpose_f64_kernel, grid, dim3(256), 0, stream, u, d, ell, scratch.ut.ptr);
        hipLaunchKernelGGL(transpose_f64_kernel, grid, dim3(256), 0, stream, v, d, ell, scratch.vt.ptr);
        MPISDF_HIP(hipGetLastError());
        scratch.u_src = u;
        scratch.v_src = v;
        scratch.ell = ell;
        scratch.d = d;
    }
    if (scratch.scale.count < (size_t)count) scratch.scale.resize(count);
    hipLaunchKernelGGL(point_scale_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream, sw, count,
                       1.0 / std::sqrt((double)ell), scratch.scale.ptr);
    MPISDF_HIP(hipGetLastError());
    scratch.timer.begin(stream);
    kr_sketch_f64_tn(scratch.ut.ptr, scratch.vt.ptr, ell, d, phi, d, count, scratch.scale.ptr, y, ldy, ynorm2,
                     scratch.part, bad, stream);
    scratch.timer.end(stream);
}

}  // namespace mpisdf
