// Column-pivoted QR with greedy max-norm pivoting, restated for the GPU.
//
// Replaces the full LAPACK/MAGMA dgeqp3 factorisation the reference runs
// (cpp/src/anchor_select_qrcp.cpp:36-90) with an algorithm that yields the
// same pivot order but whose bulk work is GEMM-shaped and never writes Y:
//
//   residual norm of column j after k pivots:
//        nu_j^2 = ||y_j||^2 - sum_{t<k} (q_t^T y_j)^2,   Q = [q_0 .. q_{k-1}]
//
// * Blocks.  After each block of new pivots the projections Q_new^T Y are
//   folded into the per-column sums by one GEMM over all columns (plus row
//   energies sum_j (q_t^T y_j)^2, the squared row norms of R).
// * Candidates.  The C columns with the largest upper bounds
//   nu_j^2 + margin * ||y_j||^2 get explicit FP64 residual vectors
//   r_c = y_c - Q Q^T y_c (classical Gram-Schmidt, two passes).  Inside the
//   block every pivot is the exact argmax of the candidate residual norms
//   (lowest column index on ties, as dgeqp3's idamax), accepted only while it
//   beats the bound of every non-candidate column; otherwise the block ends
//   and the new q's are folded.  Each accepted pivot updates all candidate
//   residuals with one rank-1 step and recomputes their norms exactly.
//
// Non-increasing residual norms make the stale bounds valid, so the pivot
// sequence equals the greedy QRCP sequence up to FP64 rounding of the norms.
#include "qrcp.hpp"

#include "kernels.hpp"
#include "util.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>

namespace mpisdf {
namespace {

struct ArgMax {
    double value;
    int64_t slot;
    int64_t column;
};

// Device-side state of the pivot loop: pivots accepted so far, the stop
// reason of the current block (0 running, 1 block ends, 2 factorisation
// done) and |R_00|.
struct PivotState {
    int64_t k;
    int32_t stop;
    int32_t pad;
    double r00;
};
constexpr int kStepsPerSync = 16;  // pivot steps queued between host reads of the state

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <class T>
__global__ void colnorm2_kernel(const T* __restrict__ y, int64_t m, int64_t n, int64_t ldy,
                                double* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int64_t col = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (col >= n) return;
    const T* c = y + col * ldy;
    double s = 0;
    for (int64_t i = lane; i < m; i += 64) { const double v = (double)c[i]; s += v * v; }
    s = wave_sum(s);
    if (lane == 0) out[col] = s;
}

// Upper bound of the residual norm^2 of every unselected column: folded value
// plus the fold's accumulated error bound plus a relative margin.
// With tie ranks the entries are laid out in rank order, so the stable sort
// puts equal keys in rank order.
__global__ void keys_kernel(const double* __restrict__ ynorm2, const double* __restrict__ proj2,
                            const double* __restrict__ err2, const uint8_t* __restrict__ selected, double margin,
                            int64_t n, const int64_t* __restrict__ tie_rank, double* __restrict__ keys,
                            int64_t* __restrict__ idx)
{
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = tie_rank ? tie_rank[j] : j;
        keys[pos] = selected[j] ? -1.0 : fmax(ynorm2[j] - proj2[j], 0.0) + err2[j] + margin * ynorm2[j];
        idx[pos] = j;
    }
}

// y_hat_j[i] = scale[j] sum_{r < planes} d_r[i] base^-(r+1) by Horner's rule
// (exact in FP64 for base 128); colp is column j's first byte.
__device__ __forceinline__ double digit_value(const int8_t* __restrict__ colp, int64_t i, double scale, int planes,
                                              double inv_base)
{
    double v = 0;
    for (int r = planes - 1; r >= 0; --r) v = (v + (double)colp[digit_offset(i, r, planes)]) * inv_base;
    return v * scale;
}

__global__ void widen_digits_kernel(const int8_t* __restrict__ yd, int64_t P, int ytype,
                                    const double* __restrict__ yscale, int64_t j0, int64_t m, int64_t nc,
                                    double* __restrict__ out)
{
    const DType t = (DType)ytype;
    const int planes = digit_planes_of(t), colbytes = digit_column_bytes(t);
    const double inv = 1.0 / digit_base_of(t);
    const int64_t total = m * nc;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i / m, r = i - j * m;
        out[i] = digit_value(yd + (j0 + j) * colbytes * P, r, yscale[j0 + j], planes, inv);
    }
}

// Candidates of the previous block whose residual is still live (not
// selected) keep it: src[c] = their slot there, else -1.
__global__ void classify_candidates_kernel(const int64_t* __restrict__ idx, int64_t C,
                                           const int32_t* __restrict__ slot_of, const double* __restrict__ rn2_prev,
                                           int32_t* __restrict__ src)
{
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < C; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t os = slot_of[idx[c]];
        src[c] = (os >= 0 && rn2_prev[os] >= 0) ? os : -1;
    }
}

// Stable partition of the candidate list, new columns (src < 0) first, in one
// workgroup of 1024 threads: cand/srcs in the new order, *nnew = new count.
__global__ __launch_bounds__(1024) void order_candidates_kernel(const int64_t* __restrict__ idx,
                                                                const int32_t* __restrict__ src, int64_t C,
                                                                int64_t* __restrict__ cand,
                                                                int32_t* __restrict__ srcs,
                                                                int64_t* __restrict__ nnew)
{
    __shared__ int scan[1024];
    __shared__ int64_t total_new;
    const int tid = threadIdx.x;
    // pass 1: number of new columns
    int cnt = 0;
    for (int64_t c = tid; c < C; c += 1024) cnt += src[c] < 0;
    scan[tid] = cnt;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (tid < s) scan[tid] += scan[tid + s];
        __syncthreads();
    }
    if (tid == 0) total_new = scan[0];
    __syncthreads();
    const int64_t nn = total_new;
    __syncthreads();
    // pass 2: chunks of 1024 in order; inclusive scan of the new flags
    int64_t base_new = 0, base_old = nn;
    for (int64_t c0 = 0; c0 < C; c0 += 1024) {
        const int64_t c = c0 + tid;
        const int f = (c < C && src[c] < 0) ? 1 : 0;
        scan[tid] = f;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const int v = tid >= off ? scan[tid - off] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        const int incl = scan[tid];
        const int chunk_new = scan[1023];
        if (c < C) {
            const int64_t pos = f ? base_new + incl - 1 : base_old + (tid - incl);
            cand[pos] = idx[c];
            srcs[pos] = src[c];
        }
        const int64_t chunk = std::min<int64_t>(1024, C - c0);
        base_new += chunk_new;
        base_old += chunk - chunk_new;
        __syncthreads();
    }
    if (tid == 0) *nnew = nn;
}

// Residual columns of the new candidate list: the kept residual of a
// persisting candidate, or the sketch column of a new one (FP64, FP32 or
// digit planes).
__global__ void fill_candidates_kernel(const void* __restrict__ y, int ytype, int64_t ldy,
                                       const double* __restrict__ yscale, int64_t m,
                                       const int64_t* __restrict__ cand, const int32_t* __restrict__ srcs, int64_t C,
                                       const double* __restrict__ rc_prev, double* __restrict__ rc, int64_t ldr)
{
    const int64_t c = blockIdx.y;
    if (c >= C) return;
    const int32_t os = srcs[c];
    const int64_t j = cand[c];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        double v;
        if (os >= 0) v = rc_prev[i + (int64_t)os * ldr];
        else if (ytype == kF64) v = static_cast<const double*>(y)[i + j * ldy];
        else if (ytype == kF32) v = (double)static_cast<const float*>(y)[i + j * ldy];
        else
            v = digit_value(static_cast<const int8_t*>(y) + j * digit_column_bytes((DType)ytype) * ldy, i, yscale[j],
                            digit_planes_of((DType)ytype), 1.0 / digit_base_of((DType)ytype));
        rc[i + c * ldr] = v;
    }
}

__global__ void set_slots_kernel(const int64_t* __restrict__ cand, int64_t C, int32_t value_base,
                                 int32_t* __restrict__ slot_of)
{
    // value_base < 0: clear (slot_of = -1); else slot_of[cand[c]] = c
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < C; c += (int64_t)gridDim.x * blockDim.x)
        slot_of[cand[c]] = value_base < 0 ? -1 : (int32_t)c;
}

__global__ void cand_norm_kernel(const double* __restrict__ rc, int64_t m, int64_t ldr, int64_t C,
                                 double* __restrict__ rn2)
{
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (c >= C) return;
    double s = 0;
    for (int64_t i = lane; i < m; i += 64) { const double v = rc[i + c * ldr]; s += v * v; }
    s = wave_sum(s);
    if (lane == 0) rn2[c] = s;
}

// Pivot of one step: the exact argmax of the candidate norms (lowest column
// on ties), then the accept / stop rules (dgeqp3's |R_ii| stop, the numerical
// rank floor, certification against the non-candidates' bound):
//   no positive norm            -> stop 2 if every non-candidate is zero, else 1;
//   |R_ii| <= stop_abs or <= stop_rel |R_00| -> stop 2;
//   value <= bound (first pivot after a fold: < bound, or forced) -> stop 1;
//   else record the pivot (piv[k], rdiag[k]) and k += 1.
__global__ __launch_bounds__(1024) void argmax_select_kernel(const double* __restrict__ rn2,
                                                             const int64_t* __restrict__ cand, int64_t C,
                                                             const int64_t* __restrict__ tie_rank,
                                                             ArgMax* __restrict__ out, PivotState* __restrict__ st,
                                                             const double* __restrict__ bound_ptr,
                                                             int64_t block_start, double stop_abs, double stop_rel,
                                                             double margin, int force, int64_t* __restrict__ piv,
                                                             double* __restrict__ rdiag)
{
    if (st->stop) return;
    // sc holds the tie rank of the best column (its index without tie ranks)
    __shared__ double sv[1024];
    __shared__ int64_t ss[1024], sc[1024];
    double best = -2.0;
    int64_t slot = -1, col = std::numeric_limits<int64_t>::max();
    for (int64_t c = threadIdx.x; c < C; c += blockDim.x) {
        const double v = rn2[c];
        const int64_t rk = tie_rank ? tie_rank[cand[c]] : cand[c];
        if (v > best || (v == best && rk < col)) { best = v; slot = c; col = rk; }
    }
    sv[threadIdx.x] = best; ss[threadIdx.x] = slot; sc[threadIdx.x] = col;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double v = sv[threadIdx.x + s];
            const int64_t cc = sc[threadIdx.x + s];
            if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && cc < sc[threadIdx.x])) {
                sv[threadIdx.x] = v; ss[threadIdx.x] = ss[threadIdx.x + s]; sc[threadIdx.x] = cc;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const ArgMax b{sv[0], ss[0], ss[0] >= 0 ? cand[ss[0]] : -1};
        *out = b;
        const double bound2 = bound_ptr ? *bound_ptr : -1.0;
        const int64_t k = st->k;
        if (b.slot < 0 || !(b.value > 0)) { st->stop = bound2 <= 0 ? 2 : 1; return; }
        const double norm = sqrt(b.value);
        if (k == 0) st->r00 = norm;
        const double r00 = k == 0 ? norm : st->r00;
        if (norm <= stop_abs || norm <= stop_rel * r00) { st->stop = 2; return; }
        const bool first = k == block_start;
        if (!(b.value > bound2 || (first && b.value >= bound2 && margin == 0.0) || (first && force))) {
            st->stop = 1;
            return;
        }
        piv[k] = b.column;
        rdiag[k] = norm;
        st->k = k + 1;
    }
}

// The new q of a pivot step, in three kernels that keep the work spread over
// the GPU (r_p the pivot's residual, already orthogonal to every earlier q up
// to rounding; the block's earlier q's are Q[:, t0 .. t0 + nt)):
//   pivot_dots:    dpart[t][s] = chunk s of Q[:, t0 + t]^T r_p;
//   pivot_vector:  q = (r_p - sum_t d_t Q[:, t0 + t]) / ||r_p||, d_t = sum_s dpart[t][s]
//                  (the second Gram-Schmidt pass against the block), with
//                  per-workgroup partial sums of q_i^2;
//   store_q:       q / ||q|| into column k of Q and of the row-major copy Qt.
// Every sum runs in a fixed order, so the result does not depend on timing.
constexpr int kDotSplit = 8;   // m-chunks per column in pivot_dots
constexpr int kVecRows = 64;   // rows per workgroup of pivot_vector

__global__ __launch_bounds__(256) void pivot_dots_kernel(const double* __restrict__ Q, int64_t m, int64_t ldq,
                                                         int64_t t0, const double* __restrict__ rc, int64_t ldr,
                                                         const ArgMax* __restrict__ am, double* __restrict__ dpart,
                                                         const PivotState* __restrict__ st)
{
    if (st->stop) return;
    __shared__ double part[4];
    const int64_t t = blockIdx.x, chunk = (m + kDotSplit - 1) / kDotSplit;
    const int64_t i0 = (int64_t)blockIdx.y * chunk, i1 = std::min<int64_t>(m, i0 + chunk);
    const double* col = Q + (t0 + t) * ldq;
    const double* r = rc + am->slot * ldr;
    double s = 0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) s += col[i] * r[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) dpart[t * kDotSplit + blockIdx.y] = (part[0] + part[1]) + (part[2] + part[3]);
}

// Rows kVecRows b .. +kVecRows per workgroup; thread (g, l) sums the terms
// t = g mod 4 of row l, the four partials are added in order.
__global__ __launch_bounds__(256) void pivot_vector_kernel(const double* __restrict__ rc, int64_t m, int64_t ldr,
                                                           const ArgMax* __restrict__ am,
                                                           const double* __restrict__ Q, int64_t ldq, int64_t t0,
                                                           int64_t nt, const double* __restrict__ dpart,
                                                           double* __restrict__ q, double* __restrict__ npart,
                                                           uint8_t* __restrict__ selected,
                                                           const PivotState* __restrict__ st)
{
    if (st->stop) return;
    constexpr int kChunk = 1024;
    __shared__ double dsh[kChunk];
    __shared__ double red[4][kVecRows];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * kVecRows + lane;
    double v = 0;
    for (int64_t c0 = 0; c0 < nt; c0 += kChunk) {
        const int64_t cn = std::min<int64_t>(kChunk, nt - c0);
        __syncthreads();
        for (int64_t t = threadIdx.x; t < cn; t += blockDim.x) {
            const double* p = dpart + (c0 + t) * kDotSplit;
            double d = 0;
#pragma unroll
            for (int c = 0; c < kDotSplit; ++c) d += p[c];
            dsh[t] = d;
        }
        __syncthreads();
        if (i < m)
            for (int64_t t = g; t < cn; t += 4) v += dsh[t] * Q[(t0 + c0 + t) * ldq + i];
    }
    red[g][lane] = v;
    __syncthreads();
    if (g == 0) {
        double qi = 0;
        if (i < m) {
            const double proj = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
            qi = (rc[am->slot * ldr + i] - proj) / sqrt(am->value);
            q[i] = qi;
        }
        const double s2 = wave_sum(qi * qi);
        if (lane == 0) npart[blockIdx.x] = s2;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) selected[am->column] = 1;
}

__global__ __launch_bounds__(256) void store_q_kernel(const double* __restrict__ q, int64_t m,
                                                      const double* __restrict__ npart, int64_t nparts,
                                                      double* __restrict__ qk, double* __restrict__ qt_k,
                                                      int64_t ldqt, const PivotState* __restrict__ st)
{
    if (st->stop) return;
    __shared__ double total;
    if (threadIdx.x < 64) {
        double s = 0;
        for (int64_t b = threadIdx.x; b < nparts; b += 64) s += npart[b];
        s = wave_sum(s);
        if (threadIdx.x == 0) total = s;
    }
    __syncthreads();
    const double inv = 1.0 / sqrt(total);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const double v = q[i] * inv;
        qk[i] = v;
        qt_k[i * ldqt] = v;
    }
}

// For every candidate c: h = q^T r_c; r_c -= h q; rn2[c] = ||r_c||^2.  The
// pivot's own slot is retired (rn2 = -1), and so is a candidate whose norm
// falls to the block's bound or below: norms only decrease and only the first
// pivot after a fold may tie the bound, so it cannot be accepted again in
// this block, and its column is not read or written again.
__global__ __launch_bounds__(256) void cand_update_kernel(double* __restrict__ rc, int64_t m, int64_t ldr,
                                                          int64_t C, const double* __restrict__ q,
                                                          double* __restrict__ rn2, const ArgMax* __restrict__ am,
                                                          const PivotState* __restrict__ st,
                                                          const double* __restrict__ bound_ptr)
{
    if (st->stop) return;
    __shared__ double part[4];
    __shared__ double hsh;
    const int64_t c = blockIdx.x;
    if (c >= C) return;
    if (c == am->slot || rn2[c] < 0) {
        if (threadIdx.x == 0) rn2[c] = -1.0;
        return;
    }
    double* col = rc + c * ldr;
    double s = 0;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) s += q[i] * col[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) hsh = part[0] + part[1] + part[2] + part[3];
    __syncthreads();
    const double h = hsh;
    double n2 = 0;
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
        const double v = col[i] - h * q[i];
        col[i] = v;
        n2 += v * v;
    }
    n2 = wave_sum(n2);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = n2;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double v = part[0] + part[1] + part[2] + part[3];
        rn2[c] = (bound_ptr && v <= *bound_ptr) ? -1.0 : v;
    }
}

// cand_update_kernel with the column held in registers (E entries per
// thread, m <= T E): one read and one write of the column instead of two
// reads and a write; q is read twice from L2.
template <int T, int E>
__global__ __launch_bounds__(T) void cand_update_reg_kernel(double* __restrict__ rc, int64_t m, int64_t ldr,
                                                              int64_t C, const double* __restrict__ q,
                                                              double* __restrict__ rn2, const ArgMax* __restrict__ am,
                                                              const PivotState* __restrict__ st,
                                                              const double* __restrict__ bound_ptr)
{
    if (st->stop) return;
    constexpr int W = T / 64;
    __shared__ double part[W];
    __shared__ double hsh;
    const int64_t c = blockIdx.x;
    if (c >= C) return;
    if (c == am->slot || rn2[c] < 0) {
        if (threadIdx.x == 0) rn2[c] = -1.0;
        return;
    }
    double* col = rc + c * ldr;
    double v[E];
    double s = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = threadIdx.x + T * e;
        v[e] = i < m ? col[i] : 0.0;
        s += i < m ? q[i] * v[e] : 0.0;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) t += part[w];
        hsh = t;
    }
    __syncthreads();
    const double h = hsh;
    double n2 = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int64_t i = threadIdx.x + T * e;
        if (i < m) {
            const double x = v[e] - h * q[i];
            col[i] = x;
            n2 += x * x;
        }
    }
    n2 = wave_sum(n2);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = n2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) x += part[w];
        rn2[c] = (bound_ptr && x <= *bound_ptr) ? -1.0 : x;
    }
}

// out[block] = this workgroup's part of sum (Y - Q W)^2 over a column chunk;
// R holds Y chunk - Q W (m x nc).  The parts are added in block order.
__global__ void sumsq_kernel(const double* __restrict__ r, int64_t count, double* __restrict__ out)
{
    __shared__ double part[4];
    double s = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        s += r[i] * r[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

template <class T>
__global__ void widen_kernel(const T* __restrict__ y, int64_t ldy, int64_t m, int64_t nc, double* __restrict__ out)
{
    const int64_t total = m * nc;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i / m, r = i - j * m;
        out[i] = (double)y[r + j * ldy];
    }
}

// Candidate update of one pivot step: the register-resident kernel (16
// entries per thread, 256 to 1024 threads) when the column fits
// (m <= 16384), else the two-pass one.
void launch_cand_update(dim3 grid, hipStream_t s, double* rc, int64_t m, int64_t ldr, int64_t C, const double* q,
                        double* rn2, const ArgMax* am, const PivotState* st, const double* bound_ptr)
{
    if (m <= 256 * 16)
        hipLaunchKernelGGL((cand_update_reg_kernel<256, 16>), grid, dim3(256), 0, s, rc, m, ldr, C, q, rn2, am, st,
                           bound_ptr);
    else if (m <= 512 * 16)
        hipLaunchKernelGGL((cand_update_reg_kernel<512, 16>), grid, dim3(512), 0, s, rc, m, ldr, C, q, rn2, am, st,
                           bound_ptr);
    else if (m <= 1024 * 16)
        hipLaunchKernelGGL((cand_update_reg_kernel<1024, 16>), grid, dim3(1024), 0, s, rc, m, ldr, C, q, rn2, am, st,
                           bound_ptr);
    else
        hipLaunchKernelGGL(cand_update_kernel, grid, dim3(256), 0, s, rc, m, ldr, C, q, rn2, am, st, bound_ptr);
}

int grid_for(int64_t work, int threads = 256, int64_t cap = 65536)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>((work + threads - 1) / threads, cap));
}

}  // namespace

QrcpResult qrcp_device(const void* y, DType ytype, int64_t m, int64_t n, int64_t ldy, const QrcpOptions& opt,
                       hipStream_t s)
{
    const bool digits = is_digit_type(ytype);
    const int planes = digit_planes_of(ytype);
    if (ytype != kF64 && ytype != kF32 && !digits) throw std::invalid_argument("QRCP input must be f64, f32 or digits");
    if (digits && (!opt.ynorm2 || !opt.yscale || !(opt.margin > 0)))
        throw std::invalid_argument("QRCP on digit planes needs device norms, scales and a positive margin");
    if (digits && (ldy % kDigitSlice || m >= (int64_t(1) << 16)))
        throw std::invalid_argument("digit planes need a pitch that is a multiple of 64 and fewer than 65536 rows");
    if (m <= 0 || n <= 0) return {};
    if (ldy < m) throw std::invalid_argument("QRCP leading dimension smaller than rows");
    const int64_t kmax = std::min({opt.kmax > 0 ? opt.kmax : std::min(m, n), m, n});
    const int64_t C0 = std::max<int64_t>(1, std::min(opt.candidates, n));
    QrcpResult res;

    DeviceBuffer<double> ynorm2(n), proj2(n), err2(n), keys(n), keys_sorted(n);
    DeviceBuffer<int64_t> idx(n), idx_sorted(n);
    DeviceBuffer<uint8_t> selected(n);
    // Q (column-major) and its row-major copy Qt; leading dimensions padded to
    // four doubles so the TN products load 32-byte vectors.
    const int64_t ldq = (m + 3) / 4 * 4, ldqt = (std::max<int64_t>(kmax, 1) + 3) / 4 * 4, ldr = ldq;
    DeviceBuffer<double> Q((size_t)ldq * std::max<int64_t>(kmax, 1)), Qt((size_t)m * ldqt), H, qv(m), dpart;
    const int64_t nvec = (m + kVecRows - 1) / kVecRows;
    DeviceBuffer<double> npart(nvec);
    // Candidate residuals, norms and columns of this block and the previous
    // one (a candidate that persists keeps its residual).
    DeviceBuffer<double> rcb[2], rn2b[2];
    DeviceBuffer<int64_t> candb[2];
    DeviceBuffer<int32_t> slot_of(n), csrc, csrcs;
    DeviceBuffer<int64_t> d_nnew(1);
    MPISDF_HIP(hipMemsetAsync(slot_of.ptr, 0xff, n * sizeof(int32_t), s));
    int cur = 0;
    int64_t Cprev = 0;
    MPISDF_HIP(hipMemsetAsync(Qt.ptr, 0, Qt.count * sizeof(double), s));
    DeviceBuffer<double> energy(std::max<int64_t>(kmax, 1)), tail(1);
    DeviceBuffer<ArgMax> am(1);
    DeviceBuffer<PivotState> pstate(1);
    DeviceBuffer<int64_t> piv_dev(std::max<int64_t>(kmax, 1));
    DeviceBuffer<double> rdiag_dev(std::max<int64_t>(kmax, 1));
    MPISDF_HIP(hipMemsetAsync(pstate.ptr, 0, sizeof(PivotState), s));
    MPISDF_HIP(hipMemsetAsync(proj2.ptr, 0, n * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(err2.ptr, 0, n * sizeof(double), s));
    MPISDF_HIP(hipMemsetAsync(selected.ptr, 0, n, s));
    MPISDF_HIP(hipMemsetAsync(energy.ptr, 0, std::max<int64_t>(kmax, 1) * sizeof(double), s));

    if (opt.ynorm2)
        MPISDF_HIP(hipMemcpyAsync(ynorm2.ptr, opt.ynorm2, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    else if (ytype == kF64)
        hipLaunchKernelGGL(colnorm2_kernel<double>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s,
                           static_cast<const double*>(y), m, n, ldy, ynorm2.ptr);
    else
        hipLaunchKernelGGL(colnorm2_kernel<float>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s,
                           static_cast<const float*>(y), m, n, ldy, ynorm2.ptr);
    MPISDF_HIP(hipGetLastError());
    {
        const auto h = ynorm2.download(s);
        double f = 0;
        for (double v : h) f += v;
        res.frob2 = f;
    }

    // hipCUB sort workspace
    size_t sort_bytes = 0;
    MPISDF_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, keys.ptr, keys_sorted.ptr,
                                                            idx.ptr, idx_sorted.ptr, (int)n, 0, 64, s));
    if (n > std::numeric_limits<int>::max()) throw std::overflow_error("QRCP column count exceeds int range");
    DeviceBuffer<uint8_t> sort_ws(std::max<size_t>(sort_bytes, 1));


    int64_t k = 0, folded = 0;
    bool done = false;
    FoldScratch fold_scratch;
    FoldI8Scratch i8_scratch;
    DeviceBuffer<double> fold_part;  // the FP64 fold's per-column-block energy partials
    const bool split = opt.split_bf16_fold && ytype == kF32;
    // Base-254 digit sketches: the pivot vectors' digits, kept for the first
    // Gram-Schmidt pass of new candidates.
    const bool digits254 = digits && digit_base_of(ytype) == 254;
    QDigitStore qdigits;
    FoldI8Rows qt_rows;         // rows of Q as digits (second pass)
    DeviceBuffer<double> gs_part;  // split-K partials of the FP64 Gram-Schmidt products
    DeviceBuffer<double> hdig;  // norms and scales of H's digit columns
    // Fold launches and each block's candidate Gram-Schmidt are bracketed by
    // HIP events; durations are read at the end.
    LaunchTimer fold_timer, gs_timer;
    // A fold streams all of Y once per group of up to 128 pivots, so a block's
    // pivots are folded in whole groups of 128 and the rest (fewer than 128)
    // are carried into the next fold: the keys then miss those projections,
    // which leaves them valid upper bounds (residual norms only decrease),
    // and certification stays exact.  The looser keys end some blocks sooner
    // ((H2O)64: 95 blocks instead of 69), but the fold, which dominates a
    // block, loses its partial launches (groups of 64 cost more blocks than
    // they save).  whole = true folds every pending pivot (row energies
    // wanted, or the previous block could certify nothing).
    const int64_t fold_group = 128;
    // Dormant columns (opt.dormant): only the prefix [0, n_active) is folded;
    // the rest keep ||y_j||^2 as their keys.  suffmax[j] = max ||y_j'||^2 over
    // j' >= j decides how far the prefix reaches.
    const bool dormant = opt.dormant && !opt.want_energy && !opt.want_tail;
    int64_t n_active = n;
    bool dormant_set = false;
    double last_bound2 = -1.0;
    std::vector<double> suffmax;
    if (dormant) {
        suffmax = ynorm2.download(s);
        suffmax.push_back(0.0);
        for (int64_t j = n; j-- > 0;) suffmax[j] = std::max(suffmax[j], suffmax[j + 1]);
    }
    // Fold pivots [q0, q0 + b) into columns [c0, c1).
    auto fold_range = [&](int64_t q0, int64_t b, int64_t c0, int64_t c1, double* energy_out) {
        const int64_t nc = c1 - c0;
        if (b <= 0 || nc <= 0) return;
        fold_timer.begin(s);
        res.fold_flops += 2.0 * (double)b * (double)m * (double)nc;
        if (digits)
            fold_projections_i8(b, nc, m, Q.ptr + q0 * ldq, ldq,
                                static_cast<const int8_t*>(y) + c0 * digit_column_bytes(ytype) * ldy, ldy, planes,
                                digit_base_of(ytype), digit_column_bytes(ytype), opt.yscale + c0, proj2.ptr + c0,
                                energy_out ? energy_out + q0 : nullptr, i8_scratch, s);
        else if (split)
            fold_projections_bf16x3(b, nc, m, Q.ptr + q0 * ldq, ldq, static_cast<const float*>(y) + c0 * ldy, ldy,
                                    proj2.ptr + c0, energy_out ? energy_out + q0 : nullptr, fold_scratch, s,
                                    ynorm2.ptr + c0, err2.ptr + c0);
        else
            fold_projections(b, nc, m, Q.ptr + q0 * ldq, ldq,
                             ytype == kF32 ? static_cast<const void*>(static_cast<const float*>(y) + c0 * ldy)
                                           : static_cast<const void*>(static_cast<const double*>(y) + c0 * ldy),
                             ytype == kF32, ldy, proj2.ptr + c0, energy_out ? energy_out + q0 : nullptr, fold_part, s);
        fold_timer.end(s);
        ++res.folds;
    };
    auto fold = [&](double* energy_out, bool whole) {
        int64_t b = k - folded;
        if (!whole && b >= fold_group) b = b / fold_group * fold_group;
        if (b <= 0) return;
        fold_range(folded, b, 0, n_active, energy_out);
        folded += b;
    };
    // Columns [n_active, c1) join the folded prefix: every pivot folded so far
    // is folded into them at once.
    auto activate = [&](int64_t c1) {
        c1 = std::min(c1, n);
        if (c1 <= n_active) return;
        fold_range(0, folded, n_active, c1, nullptr);
        n_active = c1;
    };
    // Keep the unfolded suffix below 1% of a certification bound, widening
    // the prefix once it passes 10% of the last one: the bound can fall ~8x in
    // one early block ((H2O)64), and an unfolded key above it would stall the
    // block.  A stalled block places it against its own bound.
    auto place_dormant = [&](double b2, bool now) {
        if (!dormant || b2 <= 0) return;
        if (dormant_set && !now && suffmax[n_active] <= 0.1 * b2) return;
        const double lo = 0.01 * b2;
        const int64_t want = std::lower_bound(suffmax.begin(), suffmax.end(), lo,
                                              [](double a, double v) { return a > v; }) - suffmax.begin();
        if (!dormant_set && folded == 0) n_active = std::min(want, n);  // nothing folded yet: the prefix may shrink
        else activate(want);
        dormant_set = true;
    };

    // Exact FP64 recomputation of every column's folded norm (clears the
    // split fold's error bounds); used when no pivot can be certified.
    auto refresh = [&]() {
        MPISDF_HIP(hipMemsetAsync(proj2.ptr, 0, n * sizeof(double), s));
        MPISDF_HIP(hipMemsetAsync(err2.ptr, 0, n * sizeof(double), s));
        if (k > 0)
            fold_projections(k, n_active, m, Q.ptr, ldq, y, ytype == kF32, ldy, proj2.ptr, nullptr, fold_part, s);
        folded = k;
        ++res.refreshes;
    };
    int64_t C = C0;
    bool refreshed = false, force = false;
    const int64_t Cmax = std::max<int64_t>(C0, std::min<int64_t>(n, (int64_t(1) << 30) / std::max<int64_t>(m, 1)));
    bool stalled = false;
    while (!done && k < kmax) {
        place_dormant(last_bound2, false);
        fold(opt.want_energy ? energy.ptr : nullptr, opt.want_energy || stalled);
        hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n)), dim3(256), 0, s, ynorm2.ptr, proj2.ptr, err2.ptr,
                           selected.ptr, opt.margin, n, opt.tie_rank, keys.ptr, idx.ptr);
        MPISDF_HIP(hipGetLastError());
        MPISDF_HIP(hipcub::DeviceRadixSort::SortPairsDescending(sort_ws.ptr, sort_bytes, keys.ptr, keys_sorted.ptr,
                                                                idx.ptr, idx_sorted.ptr, (int)n, 0, 64, s));
        const int64_t remaining = n - k;
        const int64_t Ceff = std::min(C, remaining);
        if (Ceff > std::numeric_limits<int32_t>::max()) throw std::overflow_error("QRCP candidate count exceeds int32");
        cur ^= 1;
        DeviceBuffer<double>& rc = rcb[cur];
        DeviceBuffer<double>& rn2 = rn2b[cur];
        DeviceBuffer<int64_t>& cand = candb[cur];
        if (rc.count < (size_t)(ldr * Ceff)) rc.resize((size_t)(ldr * Ceff));
        if (rn2.count < (size_t)Ceff) rn2.resize(Ceff);
        if (cand.count < (size_t)Ceff) cand.resize(Ceff);
        if (csrc.count < (size_t)Ceff) { csrc.resize(Ceff); csrcs.resize(Ceff); }
        double bound2 = -1.0;
        if (Ceff < remaining)
            MPISDF_HIP(hipMemcpyAsync(&bound2, keys_sorted.ptr + Ceff, sizeof(double), hipMemcpyDeviceToHost, s));
        // New candidates first, then the ones whose residual persists from the
        // previous block (already orthogonal to every q by the rank-1 updates).
        if (Cprev > 0)
            hipLaunchKernelGGL(classify_candidates_kernel, dim3(grid_for(Ceff)), dim3(256), 0, s, idx_sorted.ptr, Ceff,
                               slot_of.ptr, rn2b[cur ^ 1].ptr, csrc.ptr);
        else
            MPISDF_HIP(hipMemsetAsync(csrc.ptr, 0xff, Ceff * sizeof(int32_t), s));
        hipLaunchKernelGGL(order_candidates_kernel, dim3(1), dim3(1024), 0, s, idx_sorted.ptr, csrc.ptr, Ceff, cand.ptr,
                           csrcs.ptr, d_nnew.ptr);
        hipLaunchKernelGGL(fill_candidates_kernel, dim3(std::min(grid_for(m), 64), (unsigned)Ceff), dim3(256), 0, s, y,
                           (int)ytype, ldy, opt.yscale, m, cand.ptr, csrcs.ptr, Ceff,
                           Cprev > 0 ? rcb[cur ^ 1].ptr : nullptr, rc.ptr, ldr);
        if (Cprev > 0)
            hipLaunchKernelGGL(set_slots_kernel, dim3(grid_for(Cprev)), dim3(256), 0, s, candb[cur ^ 1].ptr, Cprev, -1,
                               slot_of.ptr);
        hipLaunchKernelGGL(set_slots_kernel, dim3(grid_for(Ceff)), dim3(256), 0, s, cand.ptr, Ceff, 0, slot_of.ptr);
        MPISDF_HIP(hipGetLastError());
        int64_t nnew = Ceff;
        MPISDF_HIP(hipMemcpyAsync(&nnew, d_nnew.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        if (bound2 > 0) last_bound2 = bound2;
        Cprev = Ceff;
        if (k > 0) {
            // Classical Gram-Schmidt, H = Q^T R, R -= Q H (via Qt): one pass over
            // every candidate, a second over the new ones (the persisting ones
            // had their two passes when they entered).
            // (digit sketches: H's columns are digitised in place for the
            // second pass, which needs a pitch of whole 64-entry slices)
            const int64_t ldh = digits254 ? (k + kDigitSlice - 1) / kDigitSlice * kDigitSlice : (k + 3) / 4 * 4;
            if (H.count < (size_t)(ldh * Ceff)) H.resize((size_t)(ldh * Ceff));
            gs_timer.begin(s);
            for (int pass = 0; pass < 2; ++pass) {
                const int64_t cols = pass == 0 ? Ceff : nnew;
                if (cols <= 0) continue;
                if (pass == 0 && digits254 && nnew > 0) {
                    // The new candidates' columns are their digit columns, so
                    // Q^T R for them runs on the INT8 digits (within ~1e-11
                    // ||y||, the truncation of q); the FP64 second pass then
                    // removes that error along Q, and R -= Q H stays FP64 (its
                    // error would fall outside span Q).  Persisting candidates
                    // carry FP64 residuals.
                    q_digits_extend(Q.ptr, ldq, m, k, kmax, qdigits, s);
                    project_candidates_i8(qdigits, m, static_cast<const int8_t*>(y), ldy, digit_column_bytes(ytype),
                                          opt.yscale, cand.ptr, nnew, H.ptr, ldh, s);
                    if (Ceff > nnew)
                        gemm_tn_f64(k, Ceff - nnew, m, 1.0, Q.ptr, ldq, rc.ptr + nnew * ldr, ldr, 0.0,
                                    H.ptr + nnew * ldh, ldh, s);
                } else {
                    gemm_tn_f64_split(k, cols, m, 1.0, Q.ptr, ldq, rc.ptr, ldr, 0.0, H.ptr, ldh, gs_part, s);
                }
                if (pass == 1 && digits254 && ldh <= kDigitRowsMaxF64) {
                    // H is the first pass's error along Q (~1e-11 ||y||), so
                    // R -= Q H needs only a few digits of relative accuracy:
                    // rows of Q and columns of H as base-254 digits.
                    split_rows_i8(Qt.ptr, ldqt, m, k, qt_rows, s);
                    if (hdig.count < (size_t)(2 * cols)) hdig.resize((size_t)(2 * cols));
                    digitize_columns_f64_max(H.ptr, k, cols, ldh, hdig.ptr, hdig.ptr + cols, s);
                    fold_rows_subtract_i8(qt_rows, reinterpret_cast<const int8_t*>(H.ptr), ldh, hdig.ptr + cols, cols,
                                          rc.ptr, ldr, s);
                } else {
                    gemm_tn_f64_split(m, cols, k, -1.0, Qt.ptr, ldqt, H.ptr, ldh, 1.0, rc.ptr, ldr, gs_part, s);
                }
                res.gs_flops += 4.0 * (double)k * (double)m * (double)cols;
            }
            gs_timer.end(s);
        }
        hipLaunchKernelGGL(cand_norm_kernel, dim3((unsigned)((Ceff + 3) / 4)), dim3(256), 0, s, rc.ptr, m, ldr, Ceff,
                           rn2.ptr);
        MPISDF_HIP(hipGetLastError());
        ++res.blocks;
        const int64_t block_start = k;
        if (dpart.count < (size_t)(kDotSplit * std::max<int64_t>(kmax, 1)))
            dpart.resize((size_t)(kDotSplit * std::max<int64_t>(kmax, 1)));
        // Pivot steps run on the device in groups of kStepsPerSync: each step's
        // argmax kernel decides (certification against the non-candidates'
        // bound: the first pivot after a fold may tie it, as candidates are
        // the lowest-index columns among equal keys) and the step's other
        // kernels return at once after a stop.  Step i of a group assumes the
        // i pivots before it were accepted, which holds for every step that
        // runs: a stop ends the block.
        const double* bound_ptr = Ceff < remaining ? keys_sorted.ptr + Ceff : nullptr;
        MPISDF_HIP(hipMemsetAsync(&pstate.ptr->stop, 0, sizeof(int32_t), s));
        bool block_over = false;
        while (!block_over && k < kmax) {
            const int64_t steps = std::min<int64_t>(kStepsPerSync, kmax - k);
            for (int64_t i = 0; i < steps; ++i) {
                const int64_t kk = k + i, nt = kk - block_start;
                hipLaunchKernelGGL(argmax_select_kernel, dim3(1), dim3(1024), 0, s, rn2.ptr, cand.ptr, Ceff,
                                   opt.tie_rank, am.ptr,
                                   pstate.ptr, bound_ptr, block_start, opt.stop_abs, opt.stop_rel, opt.margin,
                                   force ? 1 : 0, piv_dev.ptr, rdiag_dev.ptr);
                if (nt > 0)
                    hipLaunchKernelGGL(pivot_dots_kernel, dim3((unsigned)nt, kDotSplit), dim3(256), 0, s, Q.ptr, m, ldq,
                                       block_start, rc.ptr, ldr, am.ptr, dpart.ptr, pstate.ptr);
                hipLaunchKernelGGL(pivot_vector_kernel, dim3((unsigned)nvec), dim3(256), 0, s, rc.ptr, m, ldr, am.ptr,
                                   Q.ptr, ldq, block_start, nt, dpart.ptr, qv.ptr, npart.ptr, selected.ptr, pstate.ptr);
                hipLaunchKernelGGL(store_q_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, qv.ptr, m,
                                   npart.ptr, nvec, Q.ptr + kk * ldq, Qt.ptr + kk, ldqt, pstate.ptr);
                launch_cand_update(dim3((unsigned)Ceff), s, rc.ptr, m, ldr, Ceff, Q.ptr + kk * ldq, rn2.ptr, am.ptr,
                                   pstate.ptr, bound_ptr);
            }
            MPISDF_HIP(hipGetLastError());
            PivotState h{};
            MPISDF_HIP(hipMemcpyAsync(&h, pstate.ptr, sizeof(PivotState), hipMemcpyDeviceToHost, s));
            MPISDF_HIP(hipStreamSynchronize(s));
            if (h.k < k || h.k > k + steps) throw std::logic_error("QRCP pivot state out of range");
            k = h.k;
            if (h.stop) {
                block_over = true;
                if (h.stop == 2) done = true;
            }
        }
        if (k > block_start) {
            const int64_t nb = k - block_start;
            res.pivots.resize(k);
            res.rdiag.resize(k);
            MPISDF_HIP(hipMemcpyAsync(res.pivots.data() + block_start, piv_dev.ptr + block_start, nb * sizeof(int64_t),
                                      hipMemcpyDeviceToHost, s));
            MPISDF_HIP(hipMemcpyAsync(res.rdiag.data() + block_start, rdiag_dev.ptr + block_start, nb * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
            MPISDF_HIP(hipStreamSynchronize(s));
        }
        if (opt.debug)
            fprintf(stderr,
                    "[qrcp] block %lld: pivots %lld..%lld candidates %lld bound2 %.3e last_r %.3e folds %lld "
                    "folded columns %lld\n",
                    (long long)res.blocks, (long long)block_start, (long long)k, (long long)Ceff, bound2,
                    res.rdiag.empty() ? 0.0 : res.rdiag.back(), (long long)res.folds, (long long)n_active);
        if (k == kmax || done) break;
        if (k == block_start) {
            // No pivot could be certified: a non-candidate's bound ties or beats
            // every exact candidate.  Widen the candidate set; once it cannot
            // grow, make the bounds exact; if exact bounds still tie (FP64
            // rounding of ||y||^2 - ||Q^T y||^2), take the best candidate.
            ++res.stalls;
            stalled = true;
            if (folded < k) continue;  // first make every key exact to the pivots taken
            // then fold the unfolded columns near the bound (each such pass
            // folds at least one more column)
            if (bound2 > 0 && n_active < n && suffmax[n_active] > 0.01 * bound2) {
                place_dormant(bound2, true);
                continue;
            }
            if (Ceff < remaining && C < Cmax) C = std::min(2 * C, Cmax);
            else if (split && !refreshed) { refresh(); refreshed = true; }
            else if (force) break;  // even a forced pivot has no positive norm: nothing is left to factorise
            else force = true;
        } else {
            stalled = refreshed = force = false;
        }
    }
    if (opt.want_energy) {
        fold(energy.ptr, true);
        res.energy.resize(k);
        if (k) MPISDF_HIP(hipMemcpyAsync(res.energy.data(), energy.ptr, k * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    if (opt.want_tail) {
        // Exact residual after the last pivot: Y - Q Q^T Y, column chunks.
        MPISDF_HIP(hipMemsetAsync(tail.ptr, 0, sizeof(double), s));
        const int64_t nc = std::max<int64_t>(1, std::min<int64_t>(n, (1ll << 26) / std::max<int64_t>(m, 1)));
        DeviceBuffer<double> R((size_t)(m * nc)), tail_part(4096);
        DeviceBuffer<double> Wt(k > 0 ? (size_t)(k * nc) : 1);
        for (int64_t j0 = 0; j0 < n; j0 += nc) {
            const int64_t cnt = std::min(nc, n - j0);
            if (digits)
                hipLaunchKernelGGL(widen_digits_kernel, dim3(grid_for(m * cnt)), dim3(256), 0, s,
                                   static_cast<const int8_t*>(y), ldy, (int)ytype, opt.yscale, j0, m, cnt, R.ptr);
            else if (ytype == kF64)
                hipLaunchKernelGGL(widen_kernel<double>, dim3(grid_for(m * cnt)), dim3(256), 0, s,
                                   static_cast<const double*>(y) + j0 * ldy, ldy, m, cnt, R.ptr);
            else
                hipLaunchKernelGGL(widen_kernel<float>, dim3(grid_for(m * cnt)), dim3(256), 0, s,
                                   static_cast<const float*>(y) + j0 * ldy, ldy, m, cnt, R.ptr);
            if (k > 0) {
                for (int pass = 0; pass < 2; ++pass) {
                    gemm_tn_f64(k, cnt, m, 1.0, Q.ptr, ldq, R.ptr, m, 0.0, Wt.ptr, k, s);
                    gemm_f64(m, cnt, k, -1.0, MatView{Q.ptr, 1, ldq}, MatView{Wt.ptr, 1, k}, 1.0, MutView{R.ptr, 1, m},
                             s);
                }
            }
            const int blocks = grid_for(m * cnt, 256, 4096);
            hipLaunchKernelGGL(sumsq_kernel, dim3(blocks), dim3(256), 0, s, R.ptr, m * cnt, tail_part.ptr);
            MPISDF_HIP(hipGetLastError());
            launch_sum_fixed(tail_part.ptr, blocks, 0, 1, tail.ptr, s);
        }
        MPISDF_HIP(hipMemcpyAsync(&res.tail2, tail.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    }
    MPISDF_HIP(hipStreamSynchronize(s));
    res.fold_seconds = fold_timer.seconds();
    res.folded_columns = n_active;
    res.gs_seconds = gs_timer.seconds();
    return res;
}

}  // namespace mpisdf
