// Khatri-Rao pair sketch Y[s, g] = ell^{-1/2} sqrt(h_g) (u_s . z_g)(v_s . z_g)
// (cpp/src/anchor_select_randomized_id.cpp:281-309), column-major Y[g*ldy + s].
//
// FP64 form: two FP64 GEMMs (Z U, Z V) into batch scratch, then one fused
// Hadamard + weight epilogue that writes Y and counts non-finite entries
// (the reference's id_check_sketch, :310-317).
#include "kernels.hpp"
#include "sketch.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

__global__ void hadamard_kernel(const double* __restrict__ left, const double* __restrict__ right, int64_t count,
                                int64_t ell, const double* __restrict__ sw, double scaling,
                                double* __restrict__ y, int64_t ldy, unsigned long long* __restrict__ bad)
{
    const int64_t total = count * ell;
    unsigned long long nonfinite = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = i / ell, s = i - g * ell;
        const double v = scaling * sw[g] * left[i] * right[i];
        y[g * ldy + s] = v;
        nonfinite += !isfinite(v);
    }
    if (nonfinite) atomicAdd(bad, nonfinite);
}

}  // namespace

void sketch_f64(const double* phi, int64_t count, int64_t d, const double* u, const double* v, int64_t ell,
                const double* sw, double* y, int64_t ldy, SketchScratch& scratch, unsigned long long* bad,
                hipStream_t stream)
{
    if (count <= 0) return;
    const size_t need = (size_t)(count * ell);
    if (scratch.left.count < need) scratch.left.resize(need);
    if (scratch.right.count < need) scratch.right.resize(need);
    gemm_f64(count, ell, d, 1.0, MatView{phi, d, 1}, MatView{u, ell, 1}, 0.0, MutView{scratch.left.ptr, ell, 1},
             stream);
    gemm_f64(count, ell, d, 1.0, MatView{phi, d, 1}, MatView{v, ell, 1}, 0.0, MutView{scratch.right.ptr, ell, 1},
             stream);
    const int64_t blocks = std::min<int64_t>((count * ell + 255) / 256, 65536);
    hipLaunchKernelGGL(hadamard_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, scratch.left.ptr,
                       scratch.right.ptr, count, ell, sw, 1.0 / std::sqrt((double)ell), y, ldy, bad);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
