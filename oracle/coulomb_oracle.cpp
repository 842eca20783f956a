// CPU ORACLE -- test infrastructure, not product code.
//
// FP64 restatement of the reference's grid Coulomb metric
// (python/isdf.py:48-67 grid_coulomb_metric, :76-92 grid_cross_metric):
//   weighted = interp_vecs^T * w          (python/isdf.py:55, :79-80)
//   kernel_ij = 1 / |r_i - r_j|  where the distance is positive, else 0
//   kernel_ii = c_sphere |w_i|^(-1/3)     (python/isdf.py:66, :89)
//   metric = weighted_left^T (kernel weighted_right)
// evaluated in the reference's own order: the potential row
// V_i = sum_j kernel_ij weighted_right[j] first, then metric += wl_i^T V_i.
// Symmetrisation (python/isdf.py:67) is left to the caller.
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
#pragma omp parallel
    {
        std::vector<double> acc((size_t)(kl * kr), 0.0), pot((size_t)kr);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t b = 0; b < kr; ++b) pot[b] = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                double kern;
                if (j == i) {
                    kern = c_sphere * std::pow(std::fabs(w[i]), -1.0 / 3.0);
                } else {
                    const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                                 dz = xyz[3 * i + 2] - xyz[3 * j + 2];
                    const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    kern = dist > 0.0 ? 1.0 / dist : 0.0;
                }
                if (kern == 0.0) continue;
                for (int64_t b = 0; b < kr; ++b) pot[b] += kern * (right[b * n + j] * w[j]);
            }
            for (int64_t a = 0; a < kl; ++a) {
                const double la = left[a * n + i] * w[i];
                for (int64_t b = 0; b < kr; ++b) acc[a * kr + b] += la * pot[b];
            }
        }
#pragma omp critical
        for (int64_t t = 0; t < kl * kr; ++t) out[t] += acc[t];
    }
    return 0;
}
