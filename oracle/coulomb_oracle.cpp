// CPU ORACLE -- test infrastructure, not product code.
//
// FP64 restatement of the reference's grid Coulomb metric
// (python/isdf.py:48-67 grid_coulomb_metric, :76-92 grid_cross_metric):
//   weighted = interp_vecs^T * w          (python/isdf.py:55, :79-80)
//   kernel_ij = 1 / |r_i - r_j|  where the distance is positive, else 0
//   kernel_ii = c_sphere |w_i|^(-1/3)     (python/isdf.py:66, :89)
//   metric = weighted_left^T (kernel weighted_right)
// evaluated in the reference's own order: the potential row
// V_i = sum_j kernel_ij weighted_right[j] first (16 rows at a time), then
// metric += wl_i^T V_i.
// Symmetrisation (python/isdf.py:67) is left to the caller.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

extern "C" int oracle_grid_cross_metric(const double* left, int64_t kl, const double* right, int64_t kr,
                                        const double* xyz, const double* w, int64_t n, double* out)
{
    if (kl < 0 || kr < 0 || n < 0) return 1;
    const double pi = 3.14159265358979323846;
    const double c_sphere = (32.0 / 15.0) * pi * pi * std::pow(3.0 / (4.0 * pi), 5.0 / 3.0);
    for (int64_t t = 0; t < kl * kr; ++t) out[t] = 0.0;
    std::vector<double> wr((size_t)(kr * n));  // weighted_right, one row per vector
    for (int64_t b = 0; b < kr; ++b)
        for (int64_t j = 0; j < n; ++j) wr[b * n + j] = right[b * n + j] * w[j];
    constexpr int64_t IB = 16;  // kernel rows per task: each weighted_right row is read once per 16 rows
#pragma omp parallel
    {
        std::vector<double> acc((size_t)(kl * kr), 0.0), pot((size_t)(IB * kr)), krow((size_t)(IB * n));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i0 = 0; i0 < n; i0 += IB) {
            const int64_t ni = std::min(IB, n - i0);
            for (int64_t r = 0; r < ni; ++r) {
                const int64_t i = i0 + r;
                double* kr_i = krow.data() + r * n;
                for (int64_t j = 0; j < n; ++j) {
                    if (j == i) {
                        kr_i[j] = c_sphere * std::pow(std::fabs(w[i]), -1.0 / 3.0);
                    } else {
                        const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                                     dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                        kr_i[j] = dist > 0.0 ? 1.0 / dist : 0.0;
                    }
                }
            }
            for (int64_t b = 0; b < kr; ++b) {
                const double* rb = wr.data() + b * n;
                double sum[IB] = {};
                for (int64_t j = 0; j < n; ++j) {
                    const double v = rb[j];
                    for (int64_t r = 0; r < IB; ++r) sum[r] += krow[r * n + j] * v;  // rows past ni: unused
                }
                for (int64_t r = 0; r < ni; ++r) pot[r * kr + b] = sum[r];
            }
            for (int64_t a = 0; a < kl; ++a) {  // acc += wl_blk^T V_blk, acc read once per 16 rows
                double la[IB];
                for (int64_t r = 0; r < ni; ++r) la[r] = left[a * n + i0 + r] * w[i0 + r];
                double* acc_a = acc.data() + a * kr;
                for (int64_t b = 0; b < kr; ++b) {
                    double sum = acc_a[b];
                    for (int64_t r = 0; r < ni; ++r) sum += la[r] * pot[r * kr + b];
                    acc_a[b] = sum;
                }
            }
        }
#pragma omp critical
        for (int64_t t = 0; t < kl * kr; ++t) out[t] += acc[t];
    }
    return 0;
}
