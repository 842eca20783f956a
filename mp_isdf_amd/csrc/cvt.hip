// Centroidal Voronoi (CVT) anchors on the GPU: the counterpart of the
// reference's CVTAnchorSelection (cpp/src/anchor_select_cvt.cpp:17-204).
//
// The O(n k) work -- seeding distances, Lloyd assignments, snapping -- runs
// on the device.  The order-sensitive steps stay on the host in the
// reference's order: k-means++ sampling (std::mt19937_64 with
// long-double prefix sums over unnormalised masses), centroid sums in point
// order, reseeding of empty clusters at the largest residual, the shift test.
// Squared distances are formed with separately rounded products and sums,
// ((0 + dx^2) + dy^2) + dz^2 with FMA contraction off, so every comparison
// sees the same values as the host formulation and ties resolve identically.
#include "cvt.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <random>
#include <stdexcept>
#include <vector>

namespace mpisdf {
namespace {

constexpr int kCenterTile = 1024;  // centres staged in LDS per pass (24 KB)

// hipcc contracts a * b + c into an FMA by default (fast-honor-pragmas), and
// __dadd_rn / __dmul_rn are plain operators here, so contraction is switched
// off explicitly: each product and sum is rounded separately, as on the host.
__device__ __forceinline__ double dist2(double px, double py, double pz, double cx, double cy, double cz)
{
#pragma clang fp contract(off)
    const double dx = px - cx, dy = py - cy, dz = pz - cz;
    double d2 = 0.0;
    d2 += dx * dx;
    d2 += dy * dy;
    d2 += dz * dz;
    return d2;
}

__global__ __launch_bounds__(256) void seed_distance_kernel(const double* __restrict__ xyz, int64_t n, double cx,
                                                            double cy, double cz, int first,
                                                            double* __restrict__ nearest)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d2 = dist2(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], cx, cy, cz);
    nearest[i] = first ? d2 : (d2 < nearest[i] ? d2 : nearest[i]);
}

// label[i] = first k with the smallest squared distance; residual[i] = m_i * it.
__global__ __launch_bounds__(256) void lloyd_assign_kernel(const double* __restrict__ xyz, int64_t n,
                                                           const double* __restrict__ centers, int64_t rank,
                                                           const double* __restrict__ mass,
                                                           int64_t* __restrict__ label,
                                                           double* __restrict__ residual)
{
    __shared__ double sc[3 * kCenterTile];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    const double px = live ? xyz[3 * i] : 0.0, py = live ? xyz[3 * i + 1] : 0.0, pz = live ? xyz[3 * i + 2] : 0.0;
    double best = 1.7976931348623157e308;
    int64_t lab = 0;
    for (int64_t k0 = 0; k0 < rank; k0 += kCenterTile) {
        const int64_t kt = rank - k0 < kCenterTile ? rank - k0 : kCenterTile;
        __syncthreads();
        for (int64_t t = threadIdx.x; t < 3 * kt; t += 256) sc[t] = centers[3 * k0 + t];
        __syncthreads();
        for (int64_t k = 0; k < kt; ++k) {
            const double d2 = dist2(px, py, pz, sc[3 * k], sc[3 * k + 1], sc[3 * k + 2]);
            if (d2 < best) {
                best = d2;
                lab = k0 + k;
            }
        }
    }
    if (!live) return;
    label[i] = lab;
    residual[i] = mass[i] * best;
}

struct Best {
    double d2;
    int64_t i;
};

__device__ __forceinline__ bool better(const Best& a, const Best& b)
{
    return a.d2 < b.d2 || (a.d2 == b.d2 && a.i < b.i);
}

// Per-workgroup minimum of (d2, i) over unused sites.
__global__ __launch_bounds__(256) void snap_partial_kernel(const double* __restrict__ xyz, int64_t n, double cx,
                                                           double cy, double cz,
                                                           const unsigned char* __restrict__ used,
                                                           Best* __restrict__ part)
{
    __shared__ Best sb[256];
    Best b{__builtin_inf(), n};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        if (used[i]) continue;
        const Best c{dist2(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], cx, cy, cz), i};
        if (better(c, b)) b = c;
    }
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && better(sb[threadIdx.x + s], sb[threadIdx.x])) sb[threadIdx.x] = sb[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sb[0];
}

// Minimum over the partials; marks the site used and records it.
__global__ __launch_bounds__(256) void snap_final_kernel(const Best* __restrict__ part, int nparts, int64_t n,
                                                         unsigned char* __restrict__ used, int64_t* __restrict__ site)
{
    __shared__ Best sb[256];
    Best b{__builtin_inf(), n};
    for (int p = threadIdx.x; p < nparts; p += 256)
        if (better(part[p], b)) b = part[p];
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s && better(sb[threadIdx.x + s], sb[threadIdx.x])) sb[threadIdx.x] = sb[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *site = sb[0].i;
        if (sb[0].i < n) used[sb[0].i] = 1;
    }
}

}  // namespace

std::vector<int64_t> select_cvt(const double* points, const double* weights, const double* importance, int64_t npts,
                                int64_t rank, uint64_t seed, int64_t max_iter, double tol, hipStream_t s)
{
    if (npts < 0) throw std::invalid_argument("CVT points, quadrature weights and CVT weights must have equal lengths");
    if (rank <= 0 || max_iter <= 0 || !std::isfinite(tol) || tol < 0)
        throw std::invalid_argument("CVT requires positive rank and max_iter, finite nonnegative tol");
    std::vector<int64_t> active;
    std::vector<double> mass, hx;
    for (int64_t i = 0; i < npts; ++i) {
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(points[3 * i + d])) throw std::invalid_argument("CVT coordinates must be finite");
        if (!std::isfinite(weights[i]) || !std::isfinite(importance[i]) || importance[i] < 0)
            throw std::invalid_argument("CVT weights must be finite and nonnegative");
        // Quadrature weights gate eligible sites but do not scale the mass.
        const double m = weights[i] == 0 ? 0.0 : importance[i];
        if (m > 0) {
            active.push_back(i);
            mass.push_back(m);
            hx.insert(hx.end(), points + 3 * i, points + 3 * i + 3);
        }
    }
    const int64_t n = (int64_t)active.size();
    if (rank > n) throw std::invalid_argument("CVT requires at least rank positive-importance sites");
    ensure_device();
    DeviceBuffer<double> xyz, dmass, nearest(n), centers((size_t)(3 * rank)), residual(n);
    DeviceBuffer<int64_t> label(n);
    xyz.upload(hx.data(), hx.size(), s);
    dmass.upload(mass.data(), mass.size(), s);
    const unsigned blocks = (unsigned)((n + 255) / 256);

    std::mt19937_64 rng(seed);
    auto sample = [&](const std::vector<double>& scores) -> int64_t {
        long double total = 0;
        for (double v : scores) total += v;
        if (!(total > 0)) throw std::invalid_argument("CVT importance has zero total mass");
        std::uniform_real_distribution<long double> draw(0, total);
        const long double target = draw(rng);
        long double prefix = 0;
        int64_t last = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (scores[i] <= 0) continue;
            prefix += scores[i];
            last = i;
            if (target < prefix) return i;
        }
        return last;
    };
    // k-means++ seeding weighted by mass (anchor_select_cvt.cpp:37-83).
    std::vector<std::array<double, 3>> c(rank);
    const int64_t first = sample(mass);
    for (int d = 0; d < 3; ++d) c[0][d] = hx[3 * first + d];
    std::vector<double> hn(n), scores(n);
    for (int64_t k = 1; k < rank; ++k) {
        hipLaunchKernelGGL(seed_distance_kernel, dim3(blocks), dim3(256), 0, s, xyz.ptr, n, c[k - 1][0], c[k - 1][1],
                           c[k - 1][2], k == 1 ? 1 : 0, nearest.ptr);
        MPISDF_HIP(hipGetLastError());
        MPISDF_HIP(hipMemcpyAsync(hn.data(), nearest.ptr, n * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        long double total = 0;
        for (int64_t i = 0; i < n; ++i) {
            scores[i] = mass[i] * hn[i];
            if (!std::isfinite(scores[i])) throw std::invalid_argument("CVT weighted distance overflow");
            total += scores[i];
        }
        // Coincident sites: fall back to mass sampling.
        const int64_t pick = sample(total > 0 ? scores : mass);
        for (int d = 0; d < 3; ++d) c[k][d] = hx[3 * pick + d];
    }
    std::array<double, 3> lo, hi;
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = hx[d];
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], hx[3 * i + d]);
            hi[d] = std::max(hi[d], hx[3 * i + d]);
        }
    double scale2 = 0;
    for (int d = 0; d < 3; ++d) scale2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    const double scale = std::max(std::sqrt(scale2), 1.0);
    // Lloyd iterations (anchor_select_cvt.cpp:95-145).
    std::vector<double> hc((size_t)(3 * rank)), hr(n), sums((size_t)(4 * rank));
    std::vector<int64_t> hl(n);
    for (int64_t iteration = 0; iteration < max_iter; ++iteration) {
        for (int64_t k = 0; k < rank; ++k)
            for (int d = 0; d < 3; ++d) hc[3 * k + d] = c[k][d];
        MPISDF_HIP(hipMemcpyAsync(centers.ptr, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(lloyd_assign_kernel, dim3(blocks), dim3(256), 0, s, xyz.ptr, n, centers.ptr, rank,
                           dmass.ptr, label.ptr, residual.ptr);
        MPISDF_HIP(hipGetLastError());
        MPISDF_HIP(hipMemcpyAsync(hl.data(), label.ptr, n * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipMemcpyAsync(hr.data(), residual.ptr, n * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        std::fill(sums.begin(), sums.end(), 0.0);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t l = hl[i];
            sums[4 * l] += mass[i];
            for (int d = 0; d < 3; ++d) sums[4 * l + 1 + d] += mass[i] * hx[3 * i + d];
        }
        std::vector<std::array<double, 3>> next(rank);
        std::vector<int64_t> empty;
        for (int64_t k = 0; k < rank; ++k) {
            if (sums[4 * k] > 0) {
                for (int d = 0; d < 3; ++d) next[k][d] = sums[4 * k + 1 + d] / sums[4 * k];
            } else {
                empty.push_back(k);
            }
        }
        std::vector<unsigned char> reseeded(n, 0);
        for (auto k : empty) {
            int64_t site = n;
            double best = -1;
            for (int64_t i = 0; i < n; ++i)
                if (!reseeded[i] && hr[i] > best) {
                    best = hr[i];
                    site = i;
                }
            reseeded[site] = 1;
            for (int d = 0; d < 3; ++d) next[k][d] = hx[3 * site + d];
        }
        double shift2 = 0;
        for (int64_t k = 0; k < rank; ++k)
            for (int d = 0; d < 3; ++d) {
                const double delta = next[k][d] - c[k][d];
                shift2 += delta * delta;
            }
        c.swap(next);
        if (std::sqrt(shift2) <= tol * scale) break;
    }
    // Snap each centre, in order, to the nearest unused site (lowest index on
    // ties), then sort (anchor_select_cvt.cpp:148-171).
    const int nparts = (int)std::min<int64_t>(blocks, 1024);
    DeviceBuffer<Best> part(nparts);
    DeviceBuffer<unsigned char> used(n);
    DeviceBuffer<int64_t> sites(rank);
    MPISDF_HIP(hipMemsetAsync(used.ptr, 0, n, s));
    for (int64_t k = 0; k < rank; ++k) {
        hipLaunchKernelGGL(snap_partial_kernel, dim3(nparts), dim3(256), 0, s, xyz.ptr, n, c[k][0], c[k][1], c[k][2],
                           used.ptr, part.ptr);
        hipLaunchKernelGGL(snap_final_kernel, dim3(1), dim3(256), 0, s, part.ptr, nparts, n, used.ptr,
                           sites.ptr + k);
        MPISDF_HIP(hipGetLastError());
    }
    const auto hs = sites.download(s);
    std::vector<int64_t> indices(rank);
    for (int64_t k = 0; k < rank; ++k) {
        if (hs[k] < 0 || hs[k] >= n) throw std::logic_error("CVT snapping found no free site");
        indices[k] = active[hs[k]];
    }
    std::sort(indices.begin(), indices.end());
    return indices;
}

}  // namespace mpisdf
