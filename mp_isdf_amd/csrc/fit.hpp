// ISDF fit on selected anchors (see fit.hip).
#pragma once

#include "context.hpp"

namespace mpisdf {

struct FitFactor {
    int64_t k = 0, d = 0;
    DeviceBuffer<double> phi_a;  // [k][d] anchor AO values
    DeviceBuffer<double> linv;   // [k][k] row-major inverse of the Cholesky factor of G2
    DeviceBuffer<double> linvt;  // [k][k] row-major transpose of linv (upper triangular)
    double condition_estimate = 0, gram_diag_max = 0;
    // false: linv holds W = diag(lambda)^-1/2 V^T of the kept eigenpairs of G2
    // (retained_rank rows, then zero rows) in place of L^-1, a full matrix.
    bool triangular = true;
    int64_t retained_rank = 0;
};

// Anchor AOs, G2 = (Phi_a Phi_a^T)^2, Cholesky and triangular inverse.
// anchor_xyz is a device pointer [k][3].  When the Cholesky finds the anchor
// pairs numerically dependent, G2 is eigen-decomposed instead and its
// pseudo-inverse over the eigenvalues above 1e-14 of the largest is used
// (k <= 2048; std::runtime_error beyond).
void fit_factor(Context& ctx, const double* anchor_xyz, int64_t k, FitFactor& f, hipStream_t s);

// Arithmetic of the grid pass's products.  F64: S and Z = L^-1 S on FP64
// MFMA.  Int8Digits: both products on INT8 MFMA from five base-254 digits
// per entry under a power of two per row or column (within ~4.8e-13 of its
// largest entry), the 15 digit-pair products exact in INT32 and combined in
// FP64: S from the digits of phi_g and of the anchor rows, then
// ||L^-1 s_g||^2 from the digits of S's columns and of L^-1's rows, without
// forming Z; no theta (k <= 16384).  A pseudo-inverse fit (dependent
// anchors) keeps F64.
enum class FitProducts { F64, Int8Digits };

// Streams grid batches (device xyz[n][3], device w[n]) through
//   S = (Phi_a Phi_b^T)^2, Z = L^-1 S, theta = L^-T Z / w_anchor (host anchor_w[k]).
// err_out (host, 2 doubles): sum |w| (||phi||^4 - ||z||^2), sum |w| ||phi||^4.
// theta_host (optional): row-major [k][n], formed batch by batch as
// L^-T Z / w_anchor.  tsum_out (optional host [k]): sum_g theta[:, g] w_g;
// without theta_host it is L^-T (sum_g w_g z_g) / w_anchor, so the grid pass
// then runs only the S and Z products.
// gemm_stats (optional, host double[5]): summed HIP-event seconds of the grid
// pass's products (S, Z = L^-1 S, theta = L^-T Z), their algorithmic flops
// (2 k d + k (k+1) per point, + k (k+1) when theta is formed), timed
// intervals, then the seconds and launch count of the Z product alone.
void fit_stream(Context& ctx, const FitFactor& f, const double* anchor_w, const double* xyz, const double* w, int64_t n,
                int64_t batch, double* theta_host, double* err_out, double* tsum_out, hipStream_t s,
                double* gemm_stats = nullptr, FitProducts products = FitProducts::F64);

// Pair-space error of given interpolation vectors (theta_host: row-major
// [k][n] on the host), as python/isdf.py:176-191 evaluates it: per point
// ||phi_g||^4 - 2 c.s_g + c.(G2 c) with c = w_anchor o theta[:, g] and
// s_g = (Phi_a phi_g)^2.  err_out (host, 2): sum |w| err^2, sum |w| ||phi||^4.
// anchor_xyz, xyz, w: device; anchor_w: host.
void theta_pair_error(Context& ctx, const double* anchor_xyz, const double* anchor_w, int64_t k, const double* xyz,
                      const double* w, int64_t n, const double* theta_host, int64_t batch, double* err_out,
                      hipStream_t s);

// Grid-quadrature AO overlap out[nao][nao] (device) = sum_g w_g phi_g phi_g^T
// over device xyz[n][3], w[n], in batches added in order.
void quadrature_overlap(Context& ctx, const double* xyz, const double* w, int64_t n, int64_t batch, double* out,
                        hipStream_t s);

// Pair-reconstruction loss L = sum_g |w_g| ||phi_g phi_g^T - sum_k c_gk phi_k phi_k^T||_F^2
// at the fitted coefficients (= the squared l2 pair-space error) and its
// adjoint contracted with AO gradients: gao_out (host, [3][nao]) =
// sum over grid and anchor points of dL/dphi_mu(r) d phi_mu(r)/d r_c.  The
// nuclear gradient is dL/dR_A = -sum_{mu on A} gao[:, mu] (grid and anchor
// points held fixed).
void fit_loss_gradient(Context& ctx, const FitFactor& f, const double* anchor_xyz, const double* xyz,
                       const double* w, int64_t n, int64_t batch, double* loss_out, double* gao_out, hipStream_t s);

// Coulomb energy of the fitted density, the ISDF J energy
// E_J = tr(J D) / 2 = rho^T Z rho / 2 of python/isdf.py:130-151 with the grid
// metric Z = W_a grid_coulomb_metric(theta) W_a of isdf.py:48-67, :674-689,
// evaluated without theta or Z: with rho_p = phi_p^T D phi_p at the anchors
// and u_g = s_g^T G2^-1 rho (the fitted density at grid point g),
// E_J = u^T W K W u / 2, one O(n^2) potential sum.  dm is a device
// nao x nao symmetric matrix; anchor_xyz, xyz, w are device pointers (w
// nonzero).  gao_out (host [3][nao], optional) receives the adjoint contracted
// with AO gradients as for fit_loss_gradient; it needs the Cholesky factor
// (independent anchor pairs): the gradient of a pseudo-inverse fit raises.
void coulomb_energy_gradient(Context& ctx, const FitFactor& f, const double* anchor_xyz, const double* xyz,
                             const double* w, int64_t n, const double* dm, int64_t batch, double* energy_out,
                             double* gao_out, hipStream_t s);

}  // namespace mpisdf
