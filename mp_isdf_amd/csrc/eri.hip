// Integral-direct pivoted Cholesky of the AO-pair Coulomb matrix
//
//     V[(xy), (zw)] = (xy|zw) = int int phi_x(r) phi_y(r) phi_z(r') phi_w(r') / |r - r'|,
//
// over packed AO pairs x >= y (index x (x + 1) / 2 + y), and the Coulomb
// metric T V T^T through its factors: the reference's unblocked_greedy and
// unblocked_uniform metrics (cpp/src/cholesky_eri_greedy.cpp:53-175,
// cpp/include/cholesky_update.hpp:43-90, cpp/include/cholesky_pivot.hpp:16-36,
// cpp/src/coulomb_metric.cpp:20-98), which call libcint's int2e on the host.
//
// Integrals.  McMurchie-Davidson: a primitive pair of exponents a, b on
// centres A, B is a sum of Hermite Gaussians at P = (a A + b B) / p,
// p = a + b, with coefficients E^{ij}_t per axis (E^{00}_0 = 1,
// E^{i+1,j}_t = E^{ij}_{t-1} / 2p + X_PA E^{ij}_t + (t+1) E^{ij}_{t+1}), and
//     (ab|cd) = 2 pi^{5/2} / (p q sqrt(p + q)) sum_{tuv} E^{ab}_{tuv}
//               sum_{t'u'v'} (-1)^{t'+u'+v'} E^{cd}_{t'u'v'} R_{t+t', u+u', v+v'}
// with the Hermite integrals R from the Boys function F_n(alpha |PQ|^2),
// alpha = p q / (p + q).  Each pivot's column is one AO pair function
// (the ket) against every AO pair: the ket's contracted, transformed
// Hermite expansion is formed once per pivot (ket_kernel), and one wave per
// bra shell pair and contraction pair sums its primitive pairs against the
// ket primitives over the lanes, accumulating the bra's cartesian block
// (wave-reduced in a fixed order), then transforms it.  The diagonal
// (xy|xy) takes one wave per AO pair.
//
// Factorisation.  The factors stay in HBM (row-major [rank][npair]); each
// column is corrected by the earlier factors in the reference's k order,
// one thread per pair, and the next pivot is the first index of the largest
// residual (a reduction that does not depend on the order of the blocks).
// Uniform pivots are drawn on the host from std::mt19937_64 exactly as the
// reference draws them.
//
// Metric.  packed(T) (folding T[v, y, x] into pair (x, y)) times F^T and
// P P^T run on the FP64 MFMA TN product.
#include "eri.hpp"
#include "kernels.hpp"
#include "util.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>

namespace mpisdf {
namespace {

constexpr int kEriLMax = 2;       // s, p and d shells
constexpr int kHermPair = 35;     // Hermite terms of a pair function, L <= 4
constexpr int kHermAll = 165;     // R_{tuv} with t + u + v <= 8
constexpr double kTwoPi52 = 2.0 * 17.493418327624862846;  // 2 pi^(5/2)
constexpr double kHalfSqrtPi = 0.88622692545275801365;

struct KetPrim {
    double q, x, y, z;
    double h[kHermPair];
};

// Hermite index of (t, u, v): terms ordered by t + u + v, then u + v, then v.
__host__ __device__ constexpr int hidx(int t, int u, int v)
{
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int cart_base(int l) { return l * (l + 1) * (l + 2) / 6; }

// Cartesian powers in libcint order, l = 0, 1, 2 one after the other.
__constant__ int8_t kCartPow[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                                       {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}};

// F_n(t) = int_0^1 u^2n exp(-t u^2) du, n = 0..nmax: the series for F_nmax
// and downward recursion below t = 30, the erf form and upward recursion
// above (where exp(-t) no longer cancels).
__device__ void boys(int nmax, double t, double* f)
{
    if (t < 30.0) {
        double term = 1.0 / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 256; ++k) {
            term *= 2.0 * t / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        const double et = exp(-t);
        double fn = et * sum;
        f[nmax] = fn;
        for (int n = nmax - 1; n >= 0; --n) {
            fn = (2.0 * t * fn + et) / (2 * n + 1);
            f[n] = fn;
        }
    } else {
        const double st = sqrt(t), et = exp(-t), h = 0.5 / t;
        double fn = kHalfSqrtPi * erf(st) / st;
        f[0] = fn;
        for (int n = 0; n < nmax; ++n) {
            fn = ((2 * n + 1) * fn - et) * h;
            f[n + 1] = fn;
        }
    }
}

// 1-D Hermite coefficients e[i][j][t], i <= la, j <= lb (without the
// exp(-ab/p X_AB^2) factor).
__device__ inline void herm_e(int la, int lb, double o2p, double xpa, double xpb, double (&e)[3][3][5])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int t = 0; t < 5; ++t) e[i][j][t] = 0.0;
    e[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i + j == 0) continue;
            const int pi = i > 0 ? i - 1 : 0, pj = i > 0 ? j : j - 1;
            const double x = i > 0 ? xpa : xpb;
            const int top = i + j;
            for (int t = 0; t <= top; ++t) {
                double v = x * e[pi][pj][t];
                if (t > 0) v += o2p * e[pi][pj][t - 1];
                if (t + 1 < top) v += (t + 1) * e[pi][pj][t + 1];
                e[i][j][t] = v;
            }
        }
}

// R^0_{tuv}(alpha, X, Y, Z) for t + u + v <= L into r[hidx(t, u, v)], from
// R^n_{000} = (-2 alpha)^n F_n and
//     R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X R^{n+1}_{t,u,v}
// (likewise for u, v), one level n at a time in place: the entries of order
// N at level n need those of orders N - 1 and N - 2 at level n + 1, so each
// level is written from the highest order down.
__device__ void herm_r(int L, double alpha, double X, double Y, double Z, double* r)
{
    double f[9], pw[9];
    boys(L, alpha * (X * X + Y * Y + Z * Z), f);
    pw[0] = 1.0;
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
    r[0] = pw[L] * f[L];
    for (int m = L - 1; m >= 0; --m) {
        for (int N = L - m; N >= 1; --N)
            for (int uv = 0; uv <= N; ++uv)
                for (int v = 0; v <= uv; ++v) {
                    const int u = uv - v, t = N - uv;
                    double val;
                    if (t > 0)
                        val = X * r[hidx(t - 1, u, v)] + (t > 1 ? (t - 1) * r[hidx(t - 2, u, v)] : 0.0);
                    else if (u > 0)
                        val = Y * r[hidx(0, u - 1, v)] + (u > 1 ? (u - 1) * r[hidx(0, u - 2, v)] : 0.0);
                    else
                        val = Z * r[hidx(0, 0, v - 1)] + (v > 1 ? (v - 1) * r[hidx(0, 0, v - 2)] : 0.0);
                    r[hidx(t, u, v)] = val;
                }
        r[0] = pw[m] * f[m];
    }
}

// Hermite expansion, over primitive pair (i, j), of the AO pair function
// phi_x phi_y with x = (shell C, contraction ctrC, component mC) and
// y = (D, ctrD, mD): c_i c_j exp(-ab/p |C - D|^2) sum_{ab} T_x[a] T_y[b]
// E^a E^b, times (-1)^{t+u+v} for a ket.  h[nherm(lC + lD)].
__device__ void pair_hermite(const DevShell& C, int ctrC, int mC, const DevShell& D, int ctrD, int mD, int i, int j,
                             const double* __restrict__ env, const double* __restrict__ tabs, bool ket, double& p,
                             double* P, double* h)
{
    const double a = env[C.exp_off + i], b = env[D.exp_off + j];
    p = a + b;
    const double ip = 1.0 / p;
    const double A[3] = {C.x, C.y, C.z}, B[3] = {D.x, D.y, D.z};
    double e[3][3][3][5];
    double ab2 = 0;
    for (int k = 0; k < 3; ++k) {
        P[k] = (a * A[k] + b * B[k]) * ip;
        ab2 += (A[k] - B[k]) * (A[k] - B[k]);
        herm_e(C.l, D.l, 0.5 * ip, P[k] - A[k], P[k] - B[k], e[k]);
    }
    const double coef = env[C.coeff_off + ctrC * C.nprim + i] * env[D.coeff_off + ctrD * D.nprim + j] *
                        exp(-a * b * ip * ab2);
    const int L = C.l + D.l;
    for (int n = 0; n < nherm(L); ++n) h[n] = 0.0;
    const int ncC = ncart_of(C.l), ncD = ncart_of(D.l);
    const double* tC = tabs + C.tab + mC * ncC;
    const double* tD = tabs + D.tab + mD * ncD;
    for (int ca = 0; ca < ncC; ++ca) {
        const double ta = tC[ca];
        if (ta == 0.0) continue;
        const int ax = kCartPow[cart_base(C.l) + ca][0], ay = kCartPow[cart_base(C.l) + ca][1],
                  az = kCartPow[cart_base(C.l) + ca][2];
        for (int cb = 0; cb < ncD; ++cb) {
            const double w = ta * tD[cb] * coef;
            if (w == 0.0) continue;
            const int bx = kCartPow[cart_base(D.l) + cb][0], by = kCartPow[cart_base(D.l) + cb][1],
                      bz = kCartPow[cart_base(D.l) + cb][2];
            for (int t = 0; t <= ax + bx; ++t)
                for (int u = 0; u <= ay + by; ++u)
                    for (int v = 0; v <= az + bz; ++v)
                        h[hidx(t, u, v)] += w * e[0][ax][bx][t] * e[1][ay][by][u] * e[2][az][bz][v];
        }
    }
    if (ket)
        for (int N = 1; N <= L; N += 2)
            for (int n = nherm(N - 1); n < nherm(N); ++n) h[n] = -h[n];
}

// The pivot's ket pair function, one entry per primitive pair.
__global__ void ket_kernel(const DevShell* __restrict__ shells, const double* __restrict__ env,
                           const double* __restrict__ tabs, int sA, int ctrA, int mA, int sB, int ctrB, int mB,
                           KetPrim* __restrict__ out)
{
    const DevShell A = shells[sA], B = shells[sB];
    const int npp = A.nprim * B.nprim;
    for (int idx = threadIdx.x; idx < npp; idx += blockDim.x) {
        double p, P[3];
        pair_hermite(A, ctrA, mA, B, ctrB, mB, idx / B.nprim, idx % B.nprim, env, tabs, true, p, P, out[idx].h);
        out[idx].q = p;
        out[idx].x = P[0];
        out[idx].y = P[1];
        out[idx].z = P[2];
    }
}

// One wave per task (bra shells C >= D, contractions ctrC, ctrD): every
// column entry (xy|ket) with x in (C, ctrC), y in (D, ctrD), x >= y.
template <int LC, int LD>
__global__ __launch_bounds__(64) void eri_column_kernel(const int4* __restrict__ tasks, int ntask,
                                                        const DevShell* __restrict__ shells,
                                                        const double* __restrict__ env,
                                                        const double* __restrict__ tabs,
                                                        const KetPrim* __restrict__ ket, int nket, int lket,
                                                        double* __restrict__ column)
{
    constexpr int NCC = ncart_of(LC), NCD = ncart_of(LD), LB = LC + LD, NB = nherm(LB);
    const int task = blockIdx.x;
    if (task >= ntask) return;
    const int4 tk = tasks[task];
    const DevShell C = shells[tk.x], D = shells[tk.y];
    const int ctrC = tk.z, ctrD = tk.w, lane = threadIdx.x;
    const int L = LB + lket, NK = nherm(lket);
    double acc[NCC * NCD];
#pragma unroll
    for (int k = 0; k < NCC * NCD; ++k) acc[k] = 0.0;
    const double Cc[3] = {C.x, C.y, C.z}, Dc[3] = {D.x, D.y, D.z};
    const int combos = C.nprim * D.nprim * nket;
    for (int c = lane; c < combos; c += 64) {
        const int iq = c % nket, ij = c / nket, j = ij % D.nprim, i = ij / D.nprim;
        const double cc = env[C.coeff_off + ctrC * C.nprim + i] * env[D.coeff_off + ctrD * D.nprim + j];
        if (cc == 0.0) continue;
        const double a = env[C.exp_off + i], b = env[D.exp_off + j];
        const double p = a + b, ip = 1.0 / p;
        double P[3], ab2 = 0;
        double e[3][3][3][5];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            P[k] = (a * Cc[k] + b * Dc[k]) * ip;
            ab2 += (Cc[k] - Dc[k]) * (Cc[k] - Dc[k]);
            herm_e(LC, LD, 0.5 * ip, P[k] - Cc[k], P[k] - Dc[k], e[k]);
        }
        const KetPrim& kq = ket[iq];
        const double q = kq.q;
        double r[kHermAll];
        herm_r(L, p * q / (p + q), P[0] - kq.x, P[1] - kq.y, P[2] - kq.z, r);
        const double pref = kTwoPi52 / (p * q * sqrt(p + q)) * cc * exp(-a * b * ip * ab2);
        // W_tuv = sum over the ket's Hermite terms of h R_{t+t', u+u', v+v'}
        double W[NB];
        {
            int bi = 0;
            for (int nb = 0; nb <= LB; ++nb)
                for (int uvb = 0; uvb <= nb; ++uvb)
                    for (int vb = 0; vb <= uvb; ++vb, ++bi) {
                        const int ub = uvb - vb, tb = nb - uvb;
                        double s = 0;
                        int kj = 0;
                        for (int nk = 0; nk <= lket; ++nk)
                            for (int uvk = 0; uvk <= nk; ++uvk)
                                for (int vk = 0; vk <= uvk; ++vk, ++kj) {
                                    const int uk = uvk - vk, tk2 = nk - uvk;
                                    s += kq.h[kj] * r[hidx(tb + tk2, ub + uk, vb + vk)];
                                }
                        W[bi] = s;
                    }
            (void)NK;
        }
#pragma unroll
        for (int ca = 0; ca < NCC; ++ca) {
            const int ax = kCartPow[cart_base(LC) + ca][0], ay = kCartPow[cart_base(LC) + ca][1],
                      az = kCartPow[cart_base(LC) + ca][2];
#pragma unroll
            for (int cb = 0; cb < NCD; ++cb) {
                const int bx = kCartPow[cart_base(LD) + cb][0], by = kCartPow[cart_base(LD) + cb][1],
                          bz = kCartPow[cart_base(LD) + cb][2];
                double val = 0;
                for (int t = 0; t <= ax + bx; ++t)
                    for (int u = 0; u <= ay + by; ++u) {
                        const double exy = e[0][ax][bx][t] * e[1][ay][by][u];
                        for (int v = 0; v <= az + bz; ++v) val += exy * e[2][az][bz][v] * W[hidx(t, u, v)];
                    }
                acc[ca * NCD + cb] += pref * val;
            }
        }
    }
    // fixed-order butterfly: every lane ends with the same sums
#pragma unroll
    for (int k = 0; k < NCC * NCD; ++k)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off);
    const int ncompC = C.ncomp, ncompD = D.ncomp;
    for (int o = lane; o < ncompC * ncompD; o += 64) {
        const int mc = o / ncompD, md = o % ncompD;
        const double* tC = tabs + C.tab + mc * NCC;
        const double* tD = tabs + D.tab + md * NCD;
        double v = 0;
#pragma unroll
        for (int ca = 0; ca < NCC; ++ca)
#pragma unroll
            for (int cb = 0; cb < NCD; ++cb) v += tC[ca] * tD[cb] * acc[ca * NCD + cb];
        const int64_t x = C.ao_off + (int64_t)ctrC * ncompC + mc, y = D.ao_off + (int64_t)ctrD * ncompD + md;
        if (x >= y) column[x * (x + 1) / 2 + y] = v;
    }
}

__device__ inline void unpack_pair(int64_t pidx, int64_t& x, int64_t& y)
{
    x = (int64_t)((sqrt(8.0 * (double)pidx + 1.0) - 1.0) * 0.5);
    while (x * (x + 1) / 2 > pidx) --x;
    while ((x + 1) * (x + 2) / 2 <= pidx) ++x;
    y = pidx - x * (x + 1) / 2;
}

// (xy|xy) for every pair: one wave per pair, lanes over bra x ket primitive pairs.
__global__ __launch_bounds__(64) void eri_diag_kernel(const int4* __restrict__ aoinfo, int64_t npair,
                                                      const DevShell* __restrict__ shells,
                                                      const double* __restrict__ env,
                                                      const double* __restrict__ tabs, double* __restrict__ diag,
                                                      int* __restrict__ bad)
{
    const int64_t pidx = blockIdx.x;
    if (pidx >= npair) return;
    int64_t x, y;
    unpack_pair(pidx, x, y);
    const int4 ix = aoinfo[x], iy = aoinfo[y];
    const DevShell C = shells[ix.x], D = shells[iy.x];
    const int npp = C.nprim * D.nprim, L2 = C.l + D.l, L = 2 * L2;
    const int lane = threadIdx.x;
    double acc = 0;
    for (int c = lane; c < npp * npp; c += 64) {
        const int ib = c / npp, ik = c % npp;
        double p, q, P[3], Q[3], hb[kHermPair], hk[kHermPair], r[kHermAll];
        pair_hermite(C, ix.y, ix.z, D, iy.y, iy.z, ib / D.nprim, ib % D.nprim, env, tabs, false, p, P, hb);
        pair_hermite(C, ix.y, ix.z, D, iy.y, iy.z, ik / D.nprim, ik % D.nprim, env, tabs, true, q, Q, hk);
        herm_r(L, p * q / (p + q), P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], r);
        double s = 0;
        int bi = 0;
        for (int nb = 0; nb <= L2; ++nb)
            for (int uvb = 0; uvb <= nb; ++uvb)
                for (int vb = 0; vb <= uvb; ++vb, ++bi) {
                    if (hb[bi] == 0.0) continue;
                    const int ub = uvb - vb, tb = nb - uvb;
                    double w = 0;
                    int kj = 0;
                    for (int nk = 0; nk <= L2; ++nk)
                        for (int uvk = 0; uvk <= nk; ++uvk)
                            for (int vk = 0; vk <= uvk; ++vk, ++kj) {
                                const int uk = uvk - vk, tk2 = nk - uvk;
                                w += hk[kj] * r[hidx(tb + tk2, ub + uk, vb + vk)];
                            }
                    s += hb[bi] * w;
                }
        acc += kTwoPi52 / (p * q * sqrt(p + q)) * s;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        if (!isfinite(acc) || acc < -1e-12) *bad = 1;
        diag[pidx] = fmax(0.0, acc);
    }
}

// Column p of factor row `rank`: (col[p] - sum_{k<rank} F[k][p] F[k][pivot]) / denom,
// k in order; residual[p] -= value^2 (0 at the pivot, clamped at 0).
__global__ void cholesky_correct_kernel(const double* __restrict__ col, double* __restrict__ f, int64_t npair,
                                        int64_t rank, int64_t pivot, double denom, double* __restrict__ residual,
                                        int* __restrict__ bad)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npair) return;
    double v = col[p];
    for (int64_t k = 0; k < rank; ++k) v -= f[k * npair + p] * f[k * npair + pivot];
    v /= denom;
    f[rank * npair + p] = v;
    if (!isfinite(v)) *bad = 1;
    residual[p] = p == pivot ? 0.0 : fmax(0.0, residual[p] - v * v);
}

// Largest value and the first index holding it.
struct ArgMax {
    double v;
    int64_t i;
};
__device__ inline ArgMax better(ArgMax a, ArgMax b) { return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a; }

constexpr int kArgMaxBlocks = 512;
__global__ __launch_bounds__(256) void argmax_kernel(const double* __restrict__ v, int64_t n, ArgMax* __restrict__ part)
{
    __shared__ double sv[256];
    __shared__ int64_t si[256];
    ArgMax best{-1.0, n};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        best = better(best, ArgMax{v[i], i});
    sv[threadIdx.x] = best.v;
    si[threadIdx.x] = best.i;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const ArgMax o = better(ArgMax{sv[threadIdx.x], si[threadIdx.x]},
                                    ArgMax{sv[threadIdx.x + s], si[threadIdx.x + s]});
            sv[threadIdx.x] = o.v;
            si[threadIdx.x] = o.i;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = ArgMax{sv[0], si[0]};
}
__global__ void argmax_final_kernel(ArgMax* __restrict__ part, int nblocks)
{
    if (threadIdx.x != 0) return;
    ArgMax best = part[0];
    for (int b = 1; b < nblocks; ++b) best = better(best, part[b]);
    part[0] = best;
}

// packed[v][pair(x, y)] = t[v][x][y] + t[v][y][x] (x > y), t[v][x][x] on the diagonal.
__global__ void pack_transform_kernel(const double* __restrict__ t, int64_t nvec, int64_t nao, int64_t npair,
                                      double* __restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= nvec * npair) return;
    const int64_t v = idx / npair, pp = idx % npair;
    int64_t x, y;
    unpack_pair(pp, x, y);
    const double* tv = t + v * nao * nao;
    out[idx] = tv[x * nao + y] + (x != y ? tv[y * nao + x] : 0.0);
}

template <int LC, int LD>
void launch_column(const DeviceBuffer<int4>& tasks, int ntask, Context& ctx, const KetPrim* ket, int nket, int lket,
                   double* column, hipStream_t s)
{
    if (ntask <= 0) return;
    hipLaunchKernelGGL((eri_column_kernel<LC, LD>), dim3((unsigned)ntask), dim3(64), 0, s, tasks.ptr, ntask,
                       ctx.d_shells.ptr, ctx.d_env.ptr, ctx.d_tabs.ptr, ket, nket, lket, column);
    MPISDF_HIP(hipGetLastError());
}

using ColumnLauncher = void (*)(const DeviceBuffer<int4>&, int, Context&, const KetPrim*, int, int, double*,
                                hipStream_t);
const ColumnLauncher kColumnLaunchers[3][3] = {
    {launch_column<0, 0>, launch_column<0, 1>, launch_column<0, 2>},
    {launch_column<1, 0>, launch_column<1, 1>, launch_column<1, 2>},
    {launch_column<2, 0>, launch_column<2, 1>, launch_column<2, 2>},
};

int64_t pair_index(int64_t a, int64_t b) { return a * (a + 1) / 2 + b; }

ArgMax device_argmax(const double* v, int64_t n, DeviceBuffer<ArgMax>& part, hipStream_t s)
{
    const int nb = (int)std::min<int64_t>(kArgMaxBlocks, (n + 255) / 256);
    hipLaunchKernelGGL(argmax_kernel, dim3((unsigned)nb), dim3(256), 0, s, v, n, part.ptr);
    hipLaunchKernelGGL(argmax_final_kernel, dim3(1), dim3(64), 0, s, part.ptr, nb);
    MPISDF_HIP(hipGetLastError());
    ArgMax h{};
    MPISDF_HIP(hipMemcpyAsync(&h, part.ptr, sizeof(ArgMax), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
    return h;
}

}  // namespace

void eri_cholesky(Context& ctx, double tolerance, int64_t max_rank, CholeskyAlgorithm algorithm, uint64_t seed,
                  bool debug, EriCholesky& out, hipStream_t s)
{
    using clock = std::chrono::steady_clock;
    if (algorithm != CholeskyAlgorithm::UnblockedGreedy && algorithm != CholeskyAlgorithm::UnblockedUniform)
        throw std::invalid_argument("unknown Cholesky algorithm");
    if (!std::isfinite(tolerance) || tolerance < 0)
        throw std::invalid_argument("Cholesky tolerance must be finite and nonnegative");
    const int64_t nao = ctx.nao;
    if (nao > 0 && nao > (std::numeric_limits<int64_t>::max() / nao - 1) / 2)
        throw std::overflow_error("AO-pair dimension overflows size_t");
    const int64_t npair = nao * (nao + 1) / 2;
    if (npair == 0) throw std::invalid_argument("empty AO basis");
    if (max_rank < 0 || max_rank > npair) throw std::invalid_argument("max_rank exceeds AO-pair dimension");
    if (ctx.lmax > kEriLMax)
        throw std::invalid_argument("the ERI engine takes s, p and d shells (l <= 2)");
    if (npair > 0x7fffffffLL) throw std::overflow_error("too many AO pairs for the diagonal launch");
    // bra tasks by (l_C, l_D) and the AO -> (shell, contraction, component) table
    std::vector<int4> by_class[3][3];
    int nshell_pairs = 0;
    for (int sC = 0; sC < ctx.nshells; ++sC)
        for (int sD = 0; sD <= sC; ++sD) {
            ++nshell_pairs;
            const auto& C = ctx.shells[sC];
            const auto& D = ctx.shells[sD];
            for (int cC = 0; cC < C.nctr; ++cC)
                for (int cD = 0; cD < D.nctr; ++cD) {
                    if (sC == sD && cC < cD) continue;
                    by_class[C.l][D.l].push_back(make_int4(sC, sD, cC, cD));
                }
        }
    DeviceBuffer<int4> tasks[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) tasks[a][b].upload(by_class[a][b].data(), by_class[a][b].size(), s);
    std::vector<int4> aoinfo((size_t)nao);
    std::vector<int> ao_shell((size_t)nao);
    for (int sh = 0; sh < ctx.nshells; ++sh) {
        const auto& S = ctx.shells[sh];
        for (int c = 0; c < S.nctr; ++c)
            for (int m = 0; m < S.ncomp; ++m) {
                aoinfo[S.ao_off + c * S.ncomp + m] = make_int4(sh, c, m, 0);
                ao_shell[S.ao_off + c * S.ncomp + m] = sh;
            }
    }
    DeviceBuffer<int4> d_aoinfo;
    d_aoinfo.upload(aoinfo.data(), aoinfo.size(), s);
    int max_prim = 1;
    for (const auto& S : ctx.shells) max_prim = std::max(max_prim, S.nprim);
    DeviceBuffer<KetPrim> ket((size_t)max_prim * max_prim);
    DeviceBuffer<double> column((size_t)npair);
    DeviceBuffer<int> bad(1);
    DeviceBuffer<ArgMax> part(kArgMaxBlocks);
    MPISDF_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));

    const auto t_diag = clock::now();
    out.npair = npair;
    out.rank = 0;
    out.pivots.clear();
    out.residual.resize((size_t)npair);
    hipLaunchKernelGGL(eri_diag_kernel, dim3((unsigned)npair), dim3(64), 0, s, d_aoinfo.ptr, npair,
                       ctx.d_shells.ptr, ctx.d_env.ptr, ctx.d_tabs.ptr, out.residual.ptr, bad.ptr);
    MPISDF_HIP(hipGetLastError());
    int hbad = 0;
    MPISDF_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
    if (hbad) throw std::runtime_error("invalid ERI diagonal");
    ArgMax next = device_argmax(out.residual.ptr, npair, part, s);
    const double initial = next.v, cutoff = tolerance * initial;
    const int64_t cap = max_rank ? max_rank : npair;
    const char* name = algorithm == CholeskyAlgorithm::UnblockedUniform ? "unblocked_uniform" : "unblocked_greedy";
    if (debug)
        std::fprintf(stderr,
                     "[cholesky ao] algorithm=%s shells=%d pairs=%lld diagonal_quartets=%d initial_max_residual=%.17g "
                     "cutoff=%.17g diagonal_seconds=%g\n",
                     name, ctx.nshells, (long long)npair, nshell_pairs, initial, cutoff,
                     std::chrono::duration<double>(clock::now() - t_diag).count());
    std::mt19937_64 gen(seed);
    std::vector<double> hres;
    int64_t capacity = std::min<int64_t>(cap, 256);
    out.factors.resize((size_t)(capacity * npair));
    while (out.rank < cap) {
        if (next.v <= cutoff || next.v <= 0.0) break;
        int64_t pivot = next.i;
        double pres = next.v;
        if (algorithm == CholeskyAlgorithm::UnblockedUniform) {
            // uniform among residuals above the cutoff (cholesky_pivot.hpp:16-36)
            hres = out.residual.download(s);
            size_t eligible = 0;
            for (double v : hres)
                if (v > cutoff && v > 0.0) ++eligible;
            if (!eligible) throw std::logic_error("positive Cholesky residual has no eligible pivot");
            std::uniform_int_distribution<size_t> pick(0, eligible - 1);
            size_t ordinal = pick(gen);
            bool found = false;
            for (int64_t p = 0; p < npair; ++p)
                if (hres[p] > cutoff && hres[p] > 0.0 && ordinal-- == 0) {
                    pivot = p;
                    pres = hres[p];
                    found = true;
                    break;
                }
            if (!found) throw std::logic_error("uniform Cholesky pivot selection failed");
        }
        int64_t lo = 0, hi = nao;
        while (lo + 1 < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (pair_index(mid, 0) <= pivot) lo = mid; else hi = mid;
        }
        const int64_t mu = lo, nu = pivot - pair_index(mu, 0);
        const int sA = ao_shell[mu], sB = ao_shell[nu];
        const int4 iA = aoinfo[mu], iB = aoinfo[nu];
        const auto& A = ctx.shells[sA];
        const auto& B = ctx.shells[sB];
        const auto t_col = clock::now();
        hipLaunchKernelGGL(ket_kernel, dim3(1), dim3(256), 0, s, ctx.d_shells.ptr, ctx.d_env.ptr, ctx.d_tabs.ptr, sA,
                           iA.y, iA.z, sB, iB.y, iB.z, ket.ptr);
        MPISDF_HIP(hipGetLastError());
        const int nket = A.nprim * B.nprim, lket = A.l + B.l;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                kColumnLaunchers[a][b](tasks[a][b], (int)by_class[a][b].size(), ctx, ket.ptr, nket, lket, column.ptr,
                                       s);
        if (debug) MPISDF_HIP(hipStreamSynchronize(s));
        const auto t_upd = clock::now();
        if (out.rank == capacity) {  // grow the factor rows
            const int64_t grown = std::min<int64_t>(cap, capacity * 2);
            DeviceBuffer<double> bigger((size_t)(grown * npair));
            MPISDF_HIP(hipMemcpyAsync(bigger.ptr, out.factors.ptr, (size_t)(out.rank * npair) * sizeof(double),
                                      hipMemcpyDeviceToDevice, s));
            MPISDF_HIP(hipStreamSynchronize(s));
            out.factors = std::move(bigger);
            capacity = grown;
        }
        hipLaunchKernelGGL(cholesky_correct_kernel, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, column.ptr,
                           out.factors.ptr, npair, out.rank, pivot, std::sqrt(pres), out.residual.ptr, bad.ptr);
        MPISDF_HIP(hipGetLastError());
        next = device_argmax(out.residual.ptr, npair, part, s);
        MPISDF_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        if (hbad) throw std::runtime_error("non-finite Cholesky vector");
        out.pivots.push_back(pivot);
        ++out.rank;
        if (debug)
            std::fprintf(stderr,
                         "[cholesky ao] rank=%lld pivot=(%lld,%lld) shell_pair=(%d,%d) pivot_residual=%.17g "
                         "max_remaining=%.17g column_quartets=%d integral_seconds=%g update_seconds=%g\n",
                         (long long)out.rank, (long long)mu, (long long)nu, sA, sB, pres, next.v, nshell_pairs,
                         std::chrono::duration<double>(t_upd - t_col).count(),
                         std::chrono::duration<double>(clock::now() - t_upd).count());
    }
    if (debug)
        std::fprintf(stderr, "[cholesky ao] stop rank=%lld max_remaining=%.17g reason=%s\n", (long long)out.rank,
                     next.v, next.v == 0.0 ? "zero_residual" : next.v <= cutoff ? "tolerance" : "rank_limit");
}

void eri_coulomb_metric(Context& ctx, const double* transform, int64_t nvec, CholeskyAlgorithm algorithm,
                        double tolerance, int64_t max_rank, int64_t batch_size, bool debug, uint64_t seed,
                        double* out, hipStream_t s)
{
    using clock = std::chrono::steady_clock;
    const int64_t nao = ctx.nao;
    if (batch_size <= 0) throw std::invalid_argument("batch_size must be positive");
    if (nvec < 0) throw std::invalid_argument("expected transform[nvec,nao*nao]");
    if (nao && nvec > std::numeric_limits<int64_t>::max() / nao / nao)
        throw std::overflow_error("transform dimensions overflow size_t");
    for (int64_t i = 0; i < nvec * nao * nao; ++i)
        if (!std::isfinite(transform[i])) throw std::invalid_argument("transform must be finite");
    EriCholesky f;
    eri_cholesky(ctx, tolerance, max_rank, algorithm, seed, debug, f, s);
    const int64_t npair = f.npair, rank = f.rank;
    std::fill(out, out + nvec * nvec, 0.0);
    if (rank == 0 || nvec == 0) return;
    const auto t0 = clock::now();
    if (debug)
        std::fprintf(stderr, "[cholesky ao] projection start vectors=%lld pairs=%lld rank=%lld batches=%lld\n",
                     (long long)nvec, (long long)npair, (long long)rank,
                     (long long)(nvec / batch_size + (nvec % batch_size != 0)));
    DeviceBuffer<double> dt, packed((size_t)(nvec * npair)), proj((size_t)(nvec * rank)), metric((size_t)(nvec * nvec));
    dt.upload(transform, (size_t)(nvec * nao * nao), s);
    hipLaunchKernelGGL(pack_transform_kernel, dim3((unsigned)((nvec * npair + 255) / 256)), dim3(256), 0, s, dt.ptr,
                       nvec, nao, npair, packed.ptr);
    MPISDF_HIP(hipGetLastError());
    // proj[v][r] = sum_p packed[v][p] F[r][p]: TN with rank rows, column-major rank x nvec
    gemm_tn_f64(rank, nvec, npair, 1.0, f.factors.ptr, npair, packed.ptr, npair, 0.0, proj.ptr, rank, s);
    if (debug) {
        MPISDF_HIP(hipStreamSynchronize(s));
        std::fprintf(stderr, "[cholesky ao] projection_seconds=%g metric_gemm start\n",
                     std::chrono::duration<double>(clock::now() - t0).count());
    }
    const auto t1 = clock::now();
    gemm_tn_f64(nvec, nvec, rank, 1.0, proj.ptr, rank, proj.ptr, rank, 0.0, metric.ptr, nvec, s);
    MPISDF_HIP(hipMemcpyAsync(out, metric.ptr, (size_t)(nvec * nvec) * sizeof(double), hipMemcpyDeviceToHost, s));
    MPISDF_HIP(hipStreamSynchronize(s));
    if (debug)
        std::fprintf(stderr, "[cholesky ao] metric_gemm_seconds=%g\n",
                     std::chrono::duration<double>(clock::now() - t1).count());
}

}  // namespace mpisdf
