// Robust blockwise random pivoting (RBRP) on AO pair rows, with the optional
// interpolative decomposition: the counterpart of the reference's
// rbrp_ao_cpu / rbrp_ao_device (cpp/src/anchor_select_rbrp.cpp:186-318,
// cpp/src/anchor_select_rbrp_device.cpp:327-491).
//
// The weighted pair rows R (n x p, p = nao (nao + 1) / 2, off-diagonal pairs
// scaled by sqrt 2) live in HBM.  The products over all grid points run as
// FP64 MFMA GEMMs: the new block's projections R Q_b^T, the explicit residual
// refresh R - L Q^T, and the ID's L inv.  The sequential decisions run on the
// host in the reference's order: sampling with std::discrete_distribution on
// std::mt19937_64, the small p x b block's projections, its column-pivoted
// Householder QR and explicit Q, and the k x k SVD of the ID.
#include "rbrp.hpp"

#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>

namespace mpisdf {
namespace {

// R[g * p + t] = s_g phi_mu phi_nu c_{mu nu} over mu <= nu (row-major packing).
__global__ void pair_rows_kernel(const double* __restrict__ ao, const double* __restrict__ sw, int64_t n,
                                 int64_t nao, int64_t p, double* __restrict__ R)
{
    const int64_t g = blockIdx.x;
    if (g >= n) return;
    const double s = sw[g];
    const double* a = ao + g * nao;
    for (int64_t mu = threadIdx.x; mu < nao; mu += blockDim.x) {
        const int64_t base = mu * nao - mu * (mu - 1) / 2;  // packed offset of (mu, mu)
        const double am = s * a[mu];
        for (int64_t nu = mu; nu < nao; ++nu)
            R[g * p + base + (nu - mu)] = am * a[nu] * (mu == nu ? 1.0 : 1.4142135623730951);
    }
}

__global__ void row_sumsq_kernel(const double* __restrict__ E, int64_t n, int64_t p, double* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    double v = 0;
    for (int64_t i = 0; i < p; ++i) v += E[g * p + i] * E[g * p + i];
    out[g] = v;
}

double sum(const std::vector<double>& a) { return std::accumulate(a.begin(), a.end(), 0.0); }

// Column-pivoted Householder QR of the m x b column-major block v (LAPACK
// dgeqp3 conventions: pivot = largest remaining column norm, first on ties;
// reflectors H = I - tau u u^T with u_0 = 1 and beta = -sign(alpha) ||x||).
// On return v holds R in its upper triangle and the reflectors below;
// piv[j] is the 0-based original column of position j.
void householder_qrcp(std::vector<double>& v, int64_t m, int64_t b, std::vector<int64_t>& piv,
                      std::vector<double>& tau)
{
    piv.resize(b);
    tau.assign(b, 0.0);
    std::iota(piv.begin(), piv.end(), 0);
    const int64_t steps = std::min(m, b);
    for (int64_t j = 0; j < steps; ++j) {
        int64_t best = j;
        double best_norm = -1;
        for (int64_t c = j; c < b; ++c) {
            double s = 0;
            for (int64_t i = j; i < m; ++i) s += v[c * m + i] * v[c * m + i];
            if (s > best_norm) { best_norm = s; best = c; }
        }
        if (best != j) {
            for (int64_t i = 0; i < m; ++i) std::swap(v[j * m + i], v[best * m + i]);
            std::swap(piv[j], piv[best]);
        }
        double* x = v.data() + j * m;
        double xnorm = 0;
        for (int64_t i = j + 1; i < m; ++i) xnorm = std::hypot(xnorm, x[i]);
        const double alpha = x[j];
        if (xnorm == 0) { tau[j] = 0; continue; }
        const double beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
        tau[j] = (beta - alpha) / beta;
        const double scale = 1 / (alpha - beta);
        for (int64_t i = j + 1; i < m; ++i) x[i] *= scale;
        x[j] = beta;
        for (int64_t c = j + 1; c < b; ++c) {
            double* y = v.data() + c * m;
            double dot = y[j];
            for (int64_t i = j + 1; i < m; ++i) dot += x[i] * y[i];
            dot *= tau[j];
            y[j] -= dot;
            for (int64_t i = j + 1; i < m; ++i) y[i] -= dot * x[i];
        }
    }
}

// First `keep` columns of Q = H_0 ... H_{keep-1} (the reflectors of
// householder_qrcp), written column-major m x keep into q.
void explicit_q(const std::vector<double>& v, const std::vector<double>& tau, int64_t m, int64_t keep,
                std::vector<double>& q)
{
    q.assign((size_t)(m * keep), 0.0);
    for (int64_t c = 0; c < keep; ++c) q[c * m + c] = 1.0;
    for (int64_t j = keep; j-- > 0;) {
        const double* x = v.data() + j * m;
        for (int64_t c = j; c < keep; ++c) {
            double* y = q.data() + c * m;
            double dot = y[j];
            for (int64_t i = j + 1; i < m; ++i) dot += x[i] * y[i];
            dot *= tau[j];
            y[j] -= dot;
            for (int64_t i = j + 1; i < m; ++i) y[i] -= dot * x[i];
        }
    }
}

// One-sided Jacobi SVD of the k x k column-major a: a = U diag(s) V^T with s
// descending (Hestenes rotations until every column pair is orthogonal to
// working precision).
void jacobi_svd(std::vector<double> a, int64_t k, std::vector<double>& u, std::vector<double>& s,
                std::vector<double>& vmat)
{
    vmat.assign((size_t)(k * k), 0.0);
    for (int64_t i = 0; i < k; ++i) vmat[i * k + i] = 1.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int64_t p = 0; p < k; ++p)
            for (int64_t r = p + 1; r < k; ++r) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int64_t i = 0; i < k; ++i) {
                    alpha += a[p * k + i] * a[p * k + i];
                    beta += a[r * k + i] * a[r * k + i];
                    gamma += a[p * k + i] * a[r * k + i];
                }
                if (std::abs(gamma) <= eps * std::sqrt(alpha * beta) || gamma == 0) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2 * gamma);
                const double t = std::copysign(1.0, zeta) / (std::abs(zeta) + std::sqrt(1 + zeta * zeta));
                const double c = 1 / std::sqrt(1 + t * t), sn = c * t;
                for (int64_t i = 0; i < k; ++i) {
                    const double x = a[p * k + i], y = a[r * k + i];
                    a[p * k + i] = c * x - sn * y;
                    a[r * k + i] = sn * x + c * y;
                    const double vx = vmat[p * k + i], vy = vmat[r * k + i];
                    vmat[p * k + i] = c * vx - sn * vy;
                    vmat[r * k + i] = sn * vx + c * vy;
                }
            }
        if (!rotated) break;
    }
    std::vector<double> norms(k);
    for (int64_t c = 0; c < k; ++c) {
        double s2 = 0;
        for (int64_t i = 0; i < k; ++i) s2 += a[c * k + i] * a[c * k + i];
        norms[c] = std::sqrt(s2);
    }
    std::vector<int64_t> order(k);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return norms[x] > norms[y]; });
    s.resize(k);
    u.assign((size_t)(k * k), 0.0);
    std::vector<double> vs((size_t)(k * k));
    for (int64_t j = 0; j < k; ++j) {
        const int64_t c = order[j];
        s[j] = norms[c];
        for (int64_t i = 0; i < k; ++i) {
            u[j * k + i] = norms[c] > 0 ? a[c * k + i] / norms[c] : 0.0;
            vs[j * k + i] = vmat[c * k + i];
        }
    }
    vmat.swap(vs);  // column j of V (column-major)
}

}  // namespace

RbrpResult rbrp_ao(const double* ao_host, const double* ao_dev, const double* weights, int64_t n, int64_t nao,
                   const RbrpOptions& o, hipStream_t s)
{
    double tau_block = o.tau_block;
    if (n <= 0 || nao <= 0 || o.batch_size <= 0 || o.block_size <= 0 || !std::isfinite(o.tolerance) ||
        o.tolerance < 0 || !std::isfinite(o.rank_rtol) || o.rank_rtol <= 0 || !std::isfinite(o.svd_rtol) ||
        o.svd_rtol < 0 || o.max_rank < 0 || o.refresh_every < 0 ||
        (std::isfinite(tau_block) && (tau_block < 0 || tau_block >= 1)) ||
        (!std::isfinite(tau_block) && !std::isnan(tau_block)))
        throw std::invalid_argument("invalid RBRP dimensions, batch size or tolerances");
    const int64_t p = nao * (nao + 1) / 2;
    const int64_t cap = std::min({p, n, o.max_rank ? o.max_rank : std::min(n, p)});
    if (std::isnan(tau_block)) tau_block = o.block_size > 1 ? 1.0 / (double)o.block_size : 0.0;
    for (int64_t g = 0; g < n; ++g) {
        if (!std::isfinite(weights[g]) || weights[g] == 0)
            throw std::invalid_argument("RBRP requires finite nonzero weights");
        for (int64_t j = 0; j < nao; ++j)
            if (!std::isfinite(ao_host[g * nao + j])) throw std::invalid_argument("RBRP requires finite real AOs");
    }
    ensure_device();
    std::vector<double> sw(n);
    for (int64_t g = 0; g < n; ++g) sw[g] = std::sqrt(std::abs(weights[g]));
    // Pair rows on the device.
    DeviceBuffer<double> dao, dsw(n), R((size_t)(n * p)), E((size_t)(n * p)), dq((size_t)(cap * p)),
        dl((size_t)(n * std::max<int64_t>(cap, 1))), dots((size_t)(n * o.block_size)), dd(n);
    if (!ao_dev) {
        dao.upload(ao_host, (size_t)(n * nao), s);
        ao_dev = dao.ptr;
    }
    dsw.upload(sw.data(), n, s);
    hipLaunchKernelGGL(pair_rows_kernel, dim3((unsigned)n), dim3(64), 0, s, ao_dev, dsw.ptr, n, nao, p, R.ptr);
    MPISDF_HIP(hipGetLastError());
    auto row = [&](int64_t g, double* dst) {  // anchor_select_rbrp.cpp:26-32
        int64_t t = 0;
        const double sg = sw[g];
        for (int64_t mu = 0; mu < nao; ++mu)
            for (int64_t nu = mu; nu < nao; ++nu)
                dst[t++] = sg * ao_host[g * nao + mu] * ao_host[g * nao + nu] * (mu == nu ? 1.0 : std::sqrt(2.0));
    };
    std::vector<double> d(n);
    for (int64_t g = 0; g < n; ++g) {
        double ssum = 0;
        for (int64_t j = 0; j < nao; ++j) ssum += ao_host[g * nao + j] * ao_host[g * nao + j];
        d[g] = std::abs(weights[g]) * ssum * ssum;
        if (!std::isfinite(d[g])) throw std::overflow_error("non-finite RBRP AO pair norm");
    }
    const double total = sum(d), target = o.tolerance * o.tolerance * total;
    if (!std::isfinite(total)) throw std::overflow_error("non-finite RBRP total norm");
    std::vector<double> q, l;  // q: rank x p; l: n x rank (row-major)
    std::vector<unsigned char> chosen(n, 0);
    std::mt19937_64 rng(o.seed);
    RbrpResult out;
    out.status = "rank_cap";
    // ||R_g - L_g Q^T||^2 for unchosen rows (anchor_select_rbrp.cpp:62-81), on the GPU.
    auto residual = [&](int64_t rank) {
        MPISDF_HIP(hipMemcpyAsync(E.ptr, R.ptr, (size_t)(n * p) * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (rank > 0)
            gemm_f64(n, p, rank, -1.0, MatView{dl.ptr, cap, 1}, MatView{dq.ptr, p, 1}, 1.0, MutView{E.ptr, p, 1}, s);
        hipLaunchKernelGGL(row_sumsq_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, E.ptr, n, p, dd.ptr);
        MPISDF_HIP(hipGetLastError());
        std::vector<double> r = dd.download(s);
        for (int64_t g = 0; g < n; ++g)
            if (chosen[g]) r[g] = 0;
        return r;
    };
    int64_t iteration = 0;
    std::vector<double> v, raw(p), qb, tau;
    std::vector<int64_t> piv;
    while (true) {
        const int64_t old = (int64_t)out.indices.size();
        if (old && (sum(d) <= target || (o.refresh_every && iteration % o.refresh_every == 0))) d = residual(old);
        if (sum(d) <= target) { out.status = "tolerance"; break; }
        if (old >= cap) break;
        std::vector<int64_t> active;
        for (int64_t i = 0; i < n; ++i)
            if (d[i] > 0 && !chosen[i]) active.push_back(i);
        if (active.empty()) { out.status = "no_candidates"; break; }
        const int64_t b = std::min({o.block_size, cap - old, (int64_t)active.size()});
        std::vector<int64_t> sample;
        for (int64_t j = 0; j < b; ++j) {
            std::vector<double> weights_left;
            weights_left.reserve(active.size());
            for (auto id : active) weights_left.push_back(d[id]);
            std::discrete_distribution<size_t> dist(weights_left.begin(), weights_left.end());
            const auto pos = dist(rng);
            sample.push_back(active[pos]);
            active.erase(active.begin() + (std::ptrdiff_t)pos);
        }
        v.assign((size_t)(p * b), 0.0);
        double original2 = 0;
        for (int64_t j = 0; j < b; ++j) {
            row(sample[j], raw.data());
            for (int64_t i = 0; i < p; ++i) {
                v[j * p + i] = raw[i];
                original2 += raw[i] * raw[i];
            }
        }
        for (int pass = 0; pass < 2; ++pass)
            for (int64_t j = 0; j < b; ++j)
                for (int64_t t = 0; t < old; ++t) {
                    double dot = 0;
                    for (int64_t i = 0; i < p; ++i) dot += v[j * p + i] * q[t * p + i];
                    for (int64_t i = 0; i < p; ++i) v[j * p + i] -= dot * q[t * p + i];
                }
        householder_qrcp(v, p, b, piv, tau);
        std::vector<double> tails(b, 0.0);
        for (int64_t r = b; r-- > 0;) {
            tails[r] = r + 1 < b ? tails[r + 1] : 0;
            if (r < p)
                for (int64_t c = r; c < b; ++c) tails[r] += v[c * p + r] * v[c * p + r];
        }
        if (!(tails[0] > 0) || !std::isfinite(tails[0])) { out.status = "numerical_rank"; break; }
        int64_t keep = 0;
        for (int64_t i = 0; i < b; ++i)
            if (tails[i] >= tau_block * tails[0]) keep = i + 1;
        keep = std::min(keep, p);
        const double cutoff = o.rank_rtol * std::sqrt(original2);
        for (int64_t i = 0; i < keep; ++i)
            if (std::abs(v[i * p + i]) <= cutoff) { keep = i; break; }
        if (!keep) { out.status = "numerical_rank"; break; }
        explicit_q(v, tau, p, keep, qb);
        // Reorthogonalise against Q and within the block (anchor_select_rbrp.cpp:111-133).
        for (int64_t j = 0; j < keep; ++j) {
            double* col = qb.data() + j * p;
            for (int pass = 0; pass < 2; ++pass) {
                for (int64_t t = 0; t < old; ++t) {
                    double dot = 0;
                    for (int64_t i = 0; i < p; ++i) dot += q[t * p + i] * col[i];
                    for (int64_t i = 0; i < p; ++i) col[i] -= dot * q[t * p + i];
                }
                for (int64_t t = 0; t < j; ++t) {
                    double dot = 0;
                    for (int64_t i = 0; i < p; ++i) dot += qb[t * p + i] * col[i];
                    for (int64_t i = 0; i < p; ++i) col[i] -= dot * qb[t * p + i];
                }
            }
            double norm = 0;
            for (int64_t i = 0; i < p; ++i) norm = std::hypot(norm, col[i]);
            if (!(norm > 0) || !std::isfinite(norm)) throw std::runtime_error("RBRP block lost orthogonality");
            for (int64_t i = 0; i < p; ++i) col[i] /= norm;
        }
        // Projections of every row on the new directions: dots = R Q_b^T (GPU).
        MPISDF_HIP(hipMemcpyAsync(dq.ptr + old * p, qb.data(), (size_t)(keep * p) * sizeof(double),
                                  hipMemcpyHostToDevice, s));
        gemm_f64(n, keep, p, 1.0, MatView{R.ptr, p, 1}, MatView{dq.ptr + old * p, 1, p}, 0.0,
                 MutView{dots.ptr, keep, 1}, s);
        std::vector<double> hd((size_t)(n * keep));
        MPISDF_HIP(hipMemcpyAsync(hd.data(), dots.ptr, hd.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        std::vector<double> next((size_t)(n * (old + keep)));
        for (int64_t g = 0; g < n; ++g) {
            for (int64_t t = 0; t < old; ++t) next[g * (old + keep) + t] = l[g * old + t];
            double decrease = 0;
            for (int64_t t = 0; t < keep; ++t) {
                const double dot = chosen[g] ? 0.0 : hd[g * keep + t];
                next[g * (old + keep) + old + t] = dot;
                decrease += dot * dot;
            }
            d[g] = chosen[g] ? 0.0 : std::max(0.0, d[g] - decrease);
        }
        l.swap(next);
        q.insert(q.end(), qb.begin(), qb.end());
        // Device copy of the new L columns (rows at stride cap).
        std::vector<double> cols((size_t)(n * keep));
        for (int64_t g = 0; g < n; ++g)
            for (int64_t t = 0; t < keep; ++t) cols[g * keep + t] = l[g * (old + keep) + old + t];
        // One contiguous upload, then a device-side strided copy (a 2-D copy
        // straight from pageable host memory goes row by row).
        MPISDF_HIP(hipMemcpyAsync(dots.ptr, cols.data(), cols.size() * sizeof(double), hipMemcpyHostToDevice, s));
        MPISDF_HIP(hipMemcpy2DAsync(dl.ptr + old, cap * sizeof(double), dots.ptr, keep * sizeof(double),
                                    keep * sizeof(double), (size_t)n, hipMemcpyDeviceToDevice, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        for (int64_t j = 0; j < keep; ++j) {
            const int64_t id = sample[piv[j]];
            out.indices.push_back(id);
            chosen[id] = 1;
            d[id] = 0;
        }
        out.history.insert(out.history.end(), {(double)out.indices.size(), (double)b, (double)keep,
                                               total ? std::sqrt(sum(d) / total) : 0.0});
        ++iteration;
    }
    const int64_t k = (int64_t)out.indices.size();
    d = residual(k);
    out.relative_residual = total ? std::sqrt(sum(d) / total) : 0;
    if (out.relative_residual <= o.tolerance) out.status = "tolerance";
    else if (out.status == "tolerance") out.status = "roundoff_limited";
    if (!o.compute_id || !k) return out;
    // Interpolative decomposition (anchor_select_rbrp.cpp:135-184).
    std::vector<double> a((size_t)(k * k)), u, sv, vm;
    for (int64_t col = 0; col < k; ++col)
        for (int64_t r = 0; r < k; ++r) a[r + col * k] = l[out.indices[r] * k + col];
    jacobi_svd(a, k, u, sv, vm);
    out.condition = sv.back() ? sv.front() / sv.back() : std::numeric_limits<double>::infinity();
    for (double x : sv)
        if (x > o.svd_rtol * sv.front()) ++out.retained_rank;
    if (!out.retained_rank) throw std::runtime_error("RBRP anchor block has zero numerical rank");
    // inv = V diag(1/s) U^T, row-major k x k.
    std::vector<double> inv((size_t)(k * k), 0.0);
    for (int64_t i = 0; i < k; ++i)
        for (int64_t j = 0; j < k; ++j)
            for (int64_t t = 0; t < out.retained_rank; ++t) inv[i * k + j] += vm[t * k + i] * u[t * k + j] / sv[t];
    // H = L inv on the GPU (n x k), then the weight scaling per entry.
    DeviceBuffer<double> dinv, dh((size_t)(n * k));
    dinv.upload(inv.data(), inv.size(), s);
    gemm_f64(n, k, k, 1.0, MatView{dl.ptr, cap, 1}, MatView{dinv.ptr, k, 1}, 0.0, MutView{dh.ptr, k, 1}, s);
    const auto h = dh.download(s);
    out.interp_vecs.resize((size_t)(k * n));
    for (int64_t g = 0; g < n; ++g)
        for (int64_t ai = 0; ai < k; ++ai) {
            const int64_t idx = out.indices[ai];
            out.interp_vecs[ai * n + g] = h[g * k + ai] * sw[idx] / (weights[idx] * sw[g]);
        }
    for (int64_t ai = 0; ai < k; ++ai)
        for (int64_t bi = 0; bi < k; ++bi)
            out.interp_vecs[ai * n + out.indices[bi]] = ai == bi ? 1.0 / weights[out.indices[bi]] : 0.0;
    out.transform.assign((size_t)(k * nao * nao), 0.0);
    int64_t packed = 0;
    for (int64_t mu = 0; mu < nao; ++mu)
        for (int64_t nu = mu; nu < nao; ++nu, ++packed)
            for (int64_t ai = 0; ai < k; ++ai) {
                double value = 0;
                for (int64_t t = 0; t < k; ++t) value += inv[t * k + ai] * q[t * p + packed];
                const int64_t idx = out.indices[ai];
                value *= sw[idx] / weights[idx];
                if (mu != nu) value /= std::sqrt(2.0);
                out.transform[ai * nao * nao + mu * nao + nu] = value;
                out.transform[ai * nao * nao + nu * nao + mu] = value;
            }
    return out;
}

}  // namespace mpisdf
