// Grid Coulomb metric and ISDF J/K contractions (see coulomb.hip).
#pragma once

#include "common.hpp"

namespace mpisdf {

// Grid points whose kernel rows one block of the metric materialises for kr
// right-hand vectors: a multiple of 128 that gives the block's product about
// 1024 128 x 128 tiles, within a third of the free device memory and 32 GiB.
int64_t coulomb_rows_per_block(int64_t n, int64_t kr);

// metric[a * kr + b] = sum_ij left[a * ldl + i] w_i K_ij w_j right[b * ldr + j]
// over the n grid points (device xyz[n][3], w[n]; left kl x n and right
// kr x n row-major on the device), with
//   K_ij = 1 / |r_i - r_j|                  (i != j, r_i != r_j),
//   K_ij = 0                                (i != j, coincident points),
//   K_ii = c_sphere |w_i|^(-1/3),  c_sphere = (32/15) pi^2 (3 / (4 pi))^(5/3).
// symmetrize (kl == kr) replaces the result by (M + M^T) / 2.  Weights must
// be nonzero.  Rows of K are generated block by block (block_rows rows, a
// multiple of 128; 0 = coulomb_rows_per_block) into HBM and consumed by the
// FP64 MFMA TN product.
void grid_coulomb_metric(const double* left, int64_t kl, int64_t ldl, const double* right, int64_t kr, int64_t ldr,
                         const double* xyz, const double* w, int64_t n, bool symmetrize, double* metric,
                         int64_t block_rows, hipStream_t s);

// v[i] = w_i sum_j K_ij w_j u_j over the n grid points (device xyz[n][3],
// w[n], u[n], v[n]), K as above: the metric's kernel applied to one weighted
// vector by a direct sum, O(n^2) without materialising K.
void grid_coulomb_potential(const double* xyz, const double* w, const double* u, int64_t n, double* v, hipStream_t s);

// ISDF Coulomb (J) and exchange (K) matrices of a symmetric AO density dm
// (nao x nao) from anchor AO values x (k x nao row-major) and the anchor
// metric z (k x k):  P = x dm x^T,  J = x^T diag(z diag(P)) x,
// K = x^T (z o P) x, each symmetrised.  All pointers are device pointers.
void isdf_jk(const double* x, int64_t k, int64_t nao, const double* z, const double* dm, double* j, double* kmat,
             hipStream_t s);

}  // namespace mpisdf
