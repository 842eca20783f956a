// CPU ORACLE -- test infrastructure, not product code.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker or as the timed CPU baseline.
//
// FP64 restatement of the reference's anchor selection algorithms:
//   * AnchorSelection::select dispatch + multi-axis scales
//                                      cpp/src/anchor_select.cpp:79-173
//   * dense QRCP on weighted pairs     cpp/src/anchor_select_qrcp.cpp:95-225
//   * online (bounded batch) QRCP      cpp/src/anchor_select_online_qr.cpp:16-375
//   * randomized Khatri-Rao sketch     cpp/src/anchor_select_randomized_id.cpp:92-374
//
// geqp3 (LAPACK++ v2025.05.28 / vendor dgeqp3, absent from the reference tree)
// is restated from LAPACK's published dgeqp3/dlaqp2 algorithm: Businger-Golub
// max-norm pivoting (first index on ties), Householder reflectors as in
// dlarfg, and the LAWN-176 partial-norm downdate with recomputation when
// temp*(vn1/vn2)^2 <= sqrt(eps).  Pinned against scipy.linalg.qr(pivoting=True)
// (LAPACK dgeqp3) and a brute-force greedy residual check in tests/.
//
// Omega uses libstdc++'s std::mt19937_64 and std::normal_distribution<double>
// exactly as randomized_id.cpp:269-279 does.  tests/test_omega_pin.py pins it
// bit for bit against an independent Python restatement of MT19937-64 (itself
// pinned by the C++ standard's required 10000th draw, 9981545732273789042),
// of generate_canonical<double, 53> and of Marsaglia's polar method, with the
// seed mix and the u[a*ell + s] fill order.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
const char* oracle_last_error();
int oracle_axis_eval(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                     const double* env, int64_t nenv, int cart,
                     const double* xyz, int64_t n, const int32_t* axes, int naxes,
                     const double* mo_coeff, int64_t nocc, const double* scales,
                     double* out, int64_t* width_out);
int oracle_axis_scales(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                       const double* env, int64_t nenv, int cart,
                       const double* xyz, const double* weights, int64_t n,
                       const int32_t* axes, int naxes, const double* mo_coeff, int64_t nocc,
                       const double* axis_weights, double* scales_out);
}

namespace {

thread_local std::string g_sel_error;
const double EPS = std::numeric_limits<double>::epsilon();

struct Tables {
    const int32_t* atm; int natm;
    const int32_t* bas; int nbas;
    const double* env; int64_t nenv;
    int cart;
    const int32_t* axes; int naxes;
    const double* mo_coeff; int64_t nocc;
};

std::vector<double> features(const Tables& t, const double* xyz, int64_t n,
                             const std::vector<double>& scales, int64_t& width)
{
    if (oracle_axis_eval(t.atm, t.natm, t.bas, t.nbas, t.env, t.nenv, t.cart, xyz, 0, t.axes,
                         t.naxes, t.mo_coeff, t.nocc, nullptr, nullptr, &width))
        throw std::invalid_argument(oracle_last_error());
    std::vector<double> f((size_t)(n * width));
    if (n && oracle_axis_eval(t.atm, t.natm, t.bas, t.nbas, t.env, t.nenv, t.cart, xyz, n, t.axes,
                              t.naxes, t.mo_coeff, t.nocc, scales.empty() ? nullptr : scales.data(),
                              f.data(), &width))
        throw std::invalid_argument(oracle_last_error());
    return f;
}

double nrm2(int64_t n, const double* x)
{
    double scale = 0, ssq = 1;
    for (int64_t i = 0; i < n; ++i) {
        const double a = std::abs(x[i]);
        if (a != 0) {
            if (scale < a) { const double r = scale / a; ssq = 1 + ssq * r * r; scale = a; }
            else { const double r = a / scale; ssq += r * r; }
        }
    }
    return scale * std::sqrt(ssq);
}

// LAPACK dgeqp3 semantics (unblocked dlaqp2 form), column-major, in place.
// jpvt receives 1-based pivots, all columns free.
// kmax >= 0 stops after kmax Householder steps (the pivots and norms up to
// there are those of the full factorisation).
void geqp3(int64_t m, int64_t n, double* a, int64_t lda, int64_t* jpvt, double* tau, int64_t kmax = -1)
{
    const int64_t k = kmax >= 0 ? std::min({m, n, kmax}) : std::min(m, n);
    std::vector<double> vn1(n), vn2(n), work(n);
    for (int64_t j = 0; j < n; ++j) {
        jpvt[j] = j + 1;
        vn1[j] = vn2[j] = nrm2(m, a + j * lda);
    }
    const double tol3z = std::sqrt(EPS);
    for (int64_t i = 0; i < k; ++i) {
        int64_t pvt = i;
        for (int64_t j = i + 1; j < n; ++j) if (vn1[j] > vn1[pvt]) pvt = j;
        if (pvt != i) {
            std::swap_ranges(a + pvt * lda, a + pvt * lda + m, a + i * lda);
            std::swap(jpvt[pvt], jpvt[i]);
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        double* col = a + i * lda;
        // dlarfg on col[i..m)
        double t = 0;
        if (m - i > 1) {
            const double xnorm = nrm2(m - i - 1, col + i + 1);
            if (xnorm != 0) {
                const double alpha = col[i];
                const double beta = -std::copysign(std::hypot(alpha, xnorm), alpha);
                t = (beta - alpha) / beta;
                const double inv = 1.0 / (alpha - beta);
                for (int64_t r = i + 1; r < m; ++r) col[r] *= inv;
                col[i] = beta;
            }
        }
        tau[i] = t;
        if (i + 1 < n && t != 0) {
            const double aii = col[i];
            col[i] = 1.0;
#pragma omp parallel for schedule(static)
            for (int64_t j = i + 1; j < n; ++j) {
                double* c = a + j * lda;
                double w = 0;
                for (int64_t r = i; r < m; ++r) w += col[r] * c[r];
                w *= t;
                for (int64_t r = i; r < m; ++r) c[r] -= w * col[r];
            }
            col[i] = aii;
        }
#pragma omp parallel for schedule(static)
        for (int64_t j = i + 1; j < n; ++j) {
            if (vn1[j] == 0) continue;
            double temp = std::abs(a[j * lda + i]) / vn1[j];
            temp = std::max(0.0, 1.0 - temp * temp);
            const double ratio = vn1[j] / vn2[j];
            const double temp2 = temp * ratio * ratio;
            if (temp2 <= tol3z) {
                if (i + 1 < m) { vn1[j] = nrm2(m - i - 1, a + j * lda + i + 1); vn2[j] = vn1[j]; }
                else { vn1[j] = 0; vn2[j] = 0; }
            } else {
                vn1[j] *= std::sqrt(temp);
            }
        }
    }
}

struct Kept {
    std::vector<double> xyz, w, importance;
    std::vector<int64_t> original;
};

// randomized_id.cpp:137-145: zero-weight points, and for the density
// sketches zero-importance points, are not candidates.
Kept filter(const double* points, const double* weights, int64_t n, const double* importance = nullptr)
{
    Kept k;
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(weights[i])) throw std::invalid_argument("grid weights must be finite");
        if (weights[i] == 0 || (importance && importance[i] == 0)) continue;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(points[3 * i + c])) throw std::invalid_argument("grid coordinates must be finite");
        k.xyz.insert(k.xyz.end(), points + 3 * i, points + 3 * i + 3);
        k.w.push_back(weights[i]);
        k.original.push_back(i);
        if (importance) k.importance.push_back(importance[i]);
    }
    return k;
}

// cpp/src/anchor_select_qrcp.cpp:154-225
std::vector<int64_t> select_dense(const Tables& t, const double* points, const double* weights,
                                  int64_t n, double tol, int64_t max_rank,
                                  const std::vector<double>& scales)
{
    if (!std::isfinite(tol) || tol < 0) throw std::invalid_argument("invalid weights, axis list, or tolerance");
    Kept kept = filter(points, weights, n);
    int64_t width = 0;
    features(t, nullptr, 0, {}, width);
    const int64_t ng = (int64_t)kept.original.size();
    if (!ng) return {};
    const auto f = features(t, kept.xyz.data(), ng, scales, width);
    const int64_t m = width * width;
    std::vector<double> a((size_t)(m * ng));
    for (int64_t g = 0; g < ng; ++g) {
        const double sw = std::sqrt(std::abs(kept.w[g]));
        for (int64_t p = 0; p < m; ++p)
            a[g * m + p] = sw * f[g * width + p / width] * f[g * width + p % width];
    }
    for (double v : a) if (!std::isfinite(v)) throw std::overflow_error("non-finite weighted pair product");
    const int64_t cap = std::min({m, ng, max_rank ? max_rank : std::min(m, ng)});
    std::vector<int64_t> jpvt(ng);
    std::vector<double> tau(std::min(m, ng));
    geqp3(m, ng, a.data(), m, jpvt.data(), tau.data());
    const double first = cap ? std::abs(a[0]) : 0.0;
    std::vector<int64_t> out;
    for (int64_t i = 0; i < cap; ++i) {
        const double d = std::abs(a[i * m + i]);
        if (first == 0 || d <= tol * first) break;
        out.push_back(kept.original[jpvt[i] - 1]);
    }
    return out;
}

// cpp/src/anchor_select_online_qr.cpp:175-375, line by line.
std::vector<int64_t> select_online(const Tables& t, const double* points, const double* weights,
                                   int64_t npts, double tolerance, int64_t max_rank,
                                   int64_t batch_size, const std::vector<double>& scales,
                                   int64_t* stats)
{
    if (!std::isfinite(tolerance) || tolerance < 0 || batch_size <= 0)
        throw std::invalid_argument("invalid online QR weights, tolerance or batch size");
    int64_t width = 0;
    features(t, nullptr, 0, {}, width);
    std::vector<int64_t> original;
    for (int64_t i = 0; i < npts; ++i) {
        if (!std::isfinite(weights[i])) throw std::invalid_argument("grid weights must be finite");
        if (weights[i] == 0) continue;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(points[3 * i + c])) throw std::invalid_argument("grid coordinates must be finite");
        original.push_back(i);
    }
    if (original.empty()) return {};
    const int64_t m = width * width, n = (int64_t)original.size();
    const int64_t cap = std::min({m, n, max_rank ? max_rank : std::min(m, n)});

    auto columns = [&](const std::vector<int64_t>& ids) {
        std::vector<double> xyz;
        for (int64_t id : ids) xyz.insert(xyz.end(), points + 3 * id, points + 3 * id + 3);
        int64_t w = 0;
        const auto f = features(t, xyz.data(), (int64_t)ids.size(), scales, w);
        std::vector<double> a((size_t)(ids.size() * m));
        // Columns are independent: threads change nothing in any column's arithmetic.
#pragma omp parallel for schedule(static)
        for (size_t g = 0; g < ids.size(); ++g) {
            const double sw = std::sqrt(std::abs(weights[ids[g]]));
            for (int64_t p = 0; p < m; ++p)
                a[g * m + p] = sw * f[g * width + p / width] * f[g * width + p % width];
        }
        return a;
    };
    std::vector<double> reflectors, taus;
    auto transform = [&](std::vector<double>& a, int64_t count, int64_t rank) {
#pragma omp parallel for schedule(static)
        for (int64_t g = 0; g < count; ++g) {
            for (int64_t j = 0; j < rank; ++j) {
                double dot = 0;
                for (int64_t i = j; i < m; ++i) dot += reflectors[j * m + i] * a[g * m + i];
                dot *= taus[j];
                for (int64_t i = j; i < m; ++i) a[g * m + i] -= reflectors[j * m + i] * dot;
            }
        }
    };
    auto norms = [&](const std::vector<double>& a, int64_t count, int64_t first) {
        std::vector<double> r(count);
        bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
        for (int64_t g = 0; g < count; ++g) {
            double scale = 0, ssq = 1;
            for (int64_t i = first; i < m; ++i) {
                const double x = std::abs(a[g * m + i]);
                if (x != 0) {
                    if (scale < x) { const double q = scale / x; ssq = 1 + ssq * q * q; scale = x; }
                    else { const double q = x / scale; ssq += q * q; }
                }
            }
            r[g] = scale * std::sqrt(ssq);
            bad = bad || !std::isfinite(r[g]);
        }
        if (bad) throw std::overflow_error("non-finite online AO pair column");
        return r;
    };
    std::vector<double> estimated(n), reference(n);
    std::vector<unsigned char> selected(n, 0);
    std::vector<int64_t> result;
    int64_t full_scans = 0, candidate_checks = 0, reselections = 0, refreshed = 0;
    auto scan = [&](bool direct) {
        if (direct) ++full_scans;
        for (int64_t start = 0; start < n; start += std::min(batch_size, n - start)) {
            const int64_t count = std::min(batch_size, n - start);
            std::vector<int64_t> ids(original.begin() + start, original.begin() + start + count);
            auto a = columns(ids);
            if (direct) transform(a, count, (int64_t)result.size());
            const auto nr = norms(a, count, (int64_t)result.size());
            for (int64_t g = 0; g < count; ++g) {
                estimated[start + g] = selected[start + g] ? 0 : nr[g];
                reference[start + g] = estimated[start + g];
            }
        }
    };
    scan(false);
    const double first = *std::max_element(estimated.begin(), estimated.end());
    double frob = 0;
    for (double x : estimated) frob = std::hypot(frob, x);
    const double floor = 32 * EPS * frob;
    const double refresh_threshold = std::sqrt(EPS);
    const double gamma = (m * EPS < 0.5) ? m * EPS / (1 - m * EPS) : 1.0;
    auto winner = [&]() {
        int64_t p = 0;
        for (int64_t g = 1; g < n; ++g) if (estimated[g] > estimated[p]) p = g;
        return p;
    };
    while (true) {
        int64_t p = winner();
        if (estimated[p] <= tolerance * first || (int64_t)result.size() == cap || estimated[p] <= floor) {
            scan(true);
            p = winner();
            if (estimated[p] <= tolerance * first || (int64_t)result.size() == cap || estimated[p] <= floor)
                break;
        }
        std::vector<double> column;
        for (;;) {
            p = winner();
            auto a = columns({original[p]});
            transform(a, 1, (int64_t)result.size());
            const double norm = norms(a, 1, (int64_t)result.size())[0];
            ++candidate_checks;
            estimated[p] = reference[p] = norm;
            if (p == winner()) { column = a; break; }
            ++reselections;
        }
        if (estimated[p] <= floor || estimated[p] <= tolerance * first) {
            scan(true);
            if (estimated[winner()] <= floor || estimated[winner()] <= tolerance * first) break;
            continue;
        }
        const int64_t k = (int64_t)result.size();
        // make_reflector (online_qr.cpp:111-123)
        double tau = 0;
        {
            double tail = 0;
            for (int64_t i = k + 1; i < m; ++i) tail = std::hypot(tail, column[i]);
            const double alpha = column[k];
            if (tail == 0) { tau = 0; column[k] = 1; }
            else {
                const double beta = -std::copysign(std::hypot(alpha, tail), alpha);
                tau = (beta - alpha) / beta;
                const double divisor = alpha - beta;
                column[k] = 1;
                for (int64_t i = k + 1; i < m; ++i) column[i] /= divisor;
            }
        }
        for (int64_t i = 0; i < k; ++i) column[i] = 0;
        reflectors.insert(reflectors.end(), column.begin(), column.end());
        taus.push_back(tau);
        result.push_back(original[p]);
        selected[p] = 1;
        estimated[p] = reference[p] = 0;
        // q_k = H_0 ... H_k e_k
        std::vector<double> q(m, 0.0);
        q[k] = 1;
        for (int64_t rev = k + 1; rev > 0; --rev) {
            const int64_t j = rev - 1;
            double dot = 0;
            for (int64_t i = j; i < m; ++i) dot += reflectors[j * m + i] * q[i];
            const double factor = taus[j] * dot;
            for (int64_t i = j; i < m; ++i) q[i] -= factor * reflectors[j * m + i];
        }
        for (int64_t start = 0; start < n; start += std::min(batch_size, n - start)) {
            const int64_t count = std::min(batch_size, n - start);
            std::vector<int64_t> ids(original.begin() + start, original.begin() + start + count);
            auto a = columns(ids);
            std::vector<double> row(count), absrow(count);
#pragma omp parallel for schedule(static)
            for (int64_t g = 0; g < count; ++g) {
                double dot = 0, absdot = 0;
                for (int64_t i = 0; i < m; ++i) {
                    const double term = q[i] * a[g * m + i];
                    dot += term;
                    absdot += std::abs(term);
                }
                row[g] = dot;
                absrow[g] = absdot;
            }
            std::vector<int64_t> refresh;
            for (int64_t g = 0; g < count; ++g) {
                const int64_t id = start + g;
                if (selected[id]) continue;
                const double old = estimated[id], r = std::abs(row[g]);
                if (!std::isfinite(r) || !std::isfinite(absrow[g]))
                    throw std::overflow_error("non-finite online QR coefficient");
                const double ratio = old > 0 ? r / old : 0;
                const double temp = std::max(0.0, (1 + ratio) * (1 - ratio));
                const double next = old * std::sqrt(temp);
                const double ref_ratio = reference[id] > 0 ? old / reference[id] : 0;
                const double err = gamma * absrow[g];
                estimated[id] = next;
                if (old == 0 || ratio > 1 || temp * ref_ratio * ref_ratio <= refresh_threshold ||
                    (2 * r * err + err * err) > refresh_threshold * next * next)
                    refresh.push_back(original[id]);
            }
            if (!refresh.empty()) {
                refreshed += (int64_t)refresh.size();
                auto b = columns(refresh);
                transform(b, (int64_t)refresh.size(), k + 1);
                const auto nr = norms(b, (int64_t)refresh.size(), k + 1);
                for (size_t j = 0; j < refresh.size(); ++j) {
                    const auto id = std::lower_bound(original.begin() + start,
                                                     original.begin() + start + count, refresh[j]) - original.begin();
                    estimated[id] = reference[id] = nr[j];
                }
            }
        }
    }
    if (stats) { stats[0] = full_scans; stats[1] = candidate_checks; stats[2] = reselections; stats[3] = refreshed; }
    return result;
}

// cpp/src/anchor_select_randomized_id.cpp:269-279
void omega(int64_t d, int64_t ell, uint64_t seed, double* u, double* v)
{
    std::mt19937_64 gen_u(seed), gen_v(seed ^ 0x9e3779b97f4a7c15ULL);
    std::normal_distribution<double> normal_u(0, 1), normal_v(0, 1);
    for (int64_t i = 0; i < d * ell; ++i) {
        u[i] = normal_u(gen_u);
        v[i] = normal_v(gen_v);
    }
}

// cpp/src/anchor_select_randomized_id.cpp:299-309 (quadrature policy); y[g*ell+s]
void sketch(const double* phi, int64_t count, int64_t d, const double* u, const double* v,
            int64_t ell, const double* sw, double* y)
{
    const double scaling = 1 / std::sqrt((double)ell);
#pragma omp parallel
    {
        std::vector<double> left(ell), right(ell);
#pragma omp for schedule(static)
        for (int64_t g = 0; g < count; ++g) {
            std::fill(left.begin(), left.end(), 0.0);
            std::fill(right.begin(), right.end(), 0.0);
            for (int64_t a = 0; a < d; ++a) {
                const double value = phi[g * d + a];
                const double* ua = u + a * ell;
                const double* va = v + a * ell;
                for (int64_t s = 0; s < ell; ++s) { left[s] += value * ua[s]; right[s] += value * va[s]; }
            }
            for (int64_t s = 0; s < ell; ++s) y[g * ell + s] = scaling * sw[g] * left[s] * right[s];
        }
    }
}

struct RandResult {
    std::vector<int64_t> selected;
    double norm = 0;
    int64_t numerical_rank = 0, cap = 0;
};

// cpp/src/anchor_select_randomized_id.cpp:92-208 and :210-374.  policy 0:
// quadrature (column weight sqrt|w|); 1: AO density (sqrt of the point
// importance, after the AO squared-norm screening of :231-260); 2: the same
// with each AO row divided by its norm before the projections (:290-293,
// :301-304).
RandResult select_rand(const Tables& t, const double* points, const double* weights, int64_t npts,
                       double tolerance, int64_t max_rank, int64_t batch_size, int64_t sketch_size,
                       int64_t oversampling, uint64_t seed, const std::vector<double>& scales, int policy = 0,
                       const double* importance = nullptr, double screening = 1e-24)
{
    const bool density = policy != 0, normalized = policy == 2;
    if (density) {
        if (!importance) throw std::invalid_argument("point_importance must have shape (ngrid,)");
        for (int64_t i = 0; i < npts; ++i)
            if (!std::isfinite(importance[i]) || importance[i] < 0)
                throw std::invalid_argument("point_importance must be finite and nonnegative");
        if (!std::isfinite(screening) || screening < 0)
            throw std::invalid_argument("AO screening threshold must be finite and nonnegative");
        if (t.naxes != 1 || t.axes[0] != 0)
            throw std::invalid_argument("density Khatri-Rao sketches with non-AO axis features are NYI");
    }
    if (!std::isfinite(tolerance) || tolerance < 0 || batch_size <= 0 || oversampling <= 0 ||
        (!max_rank && sketch_size <= oversampling))
        throw std::invalid_argument("invalid randomized ID grid, tolerance, batch, sketch or rank/oversampling");
    int64_t d = 0;
    features(t, nullptr, 0, {}, d);
    const int64_t requested = max_rank ? max_rank + oversampling : sketch_size;
    Kept kept = filter(points, weights, npts, density ? importance : nullptr);
    int64_t n = (int64_t)kept.original.size();
    RandResult res;
    if (n == 0) {
        if (density && max_rank) throw std::invalid_argument("requested anchor_rank exceeds eligible candidates");
        return res;
    }
    const int64_t physical_rank = d * (d + 1) / 2;
    const int64_t backend_limit = std::numeric_limits<int64_t>::max() / n;
    const int64_t ell = requested <= backend_limit ? requested
        : std::min({requested, physical_rank + oversampling, backend_limit});
    if (ell <= oversampling || (max_rank && std::min(max_rank, physical_rank) > ell - oversampling))
        throw std::invalid_argument("randomized ID sketch/rank exceeds the ILP64 QRCP backend limit");
    std::vector<double> ao_norms;
    if (density) {
        Kept screened;
        for (int64_t start = 0; start < n; start += std::min(batch_size, n - start)) {
            const int64_t count = std::min(batch_size, n - start);
            int64_t w = 0;
            const auto phi = features(t, kept.xyz.data() + 3 * start, count, scales, w);
            for (int64_t g = 0; g < count; ++g) {
                double q = 0;
                for (int64_t a = 0; a < d; ++a) q += phi[g * d + a] * phi[g * d + a];
                if (!std::isfinite(q)) throw std::overflow_error("non-finite AO squared norm");
                if (q <= screening) continue;
                const int64_t id = start + g;
                screened.xyz.insert(screened.xyz.end(), kept.xyz.begin() + 3 * id, kept.xyz.begin() + 3 * id + 3);
                screened.w.push_back(kept.w[id]);
                screened.importance.push_back(kept.importance[id]);
                screened.original.push_back(kept.original[id]);
                ao_norms.push_back(q);
            }
        }
        kept = std::move(screened);
        n = (int64_t)kept.original.size();
    }
    if (max_rank > std::min(n, physical_rank))
        throw std::invalid_argument("requested anchor_rank exceeds eligible candidates or physical pair rank");
    if (n == 0) return res;
    std::vector<double> u((size_t)(d * ell)), v((size_t)(d * ell));
    omega(d, ell, seed, u.data(), v.data());
    std::vector<double> y((size_t)(ell * n));
    for (int64_t start = 0; start < n; start += std::min(batch_size, n - start)) {
        const int64_t count = std::min(batch_size, n - start);
        int64_t w = 0;
        auto phi = features(t, kept.xyz.data() + 3 * start, count, scales, w);
        std::vector<double> sw(count);
        for (int64_t g = 0; g < count; ++g) {
            sw[g] = std::sqrt(density ? kept.importance[start + g] : std::abs(kept.w[start + g]));
            if (normalized) {
                const double inv = 1 / std::sqrt(ao_norms[start + g]);
                for (int64_t a = 0; a < d; ++a) phi[g * d + a] *= inv;
            }
        }
        sketch(phi.data(), count, d, u.data(), v.data(), ell, sw.data(), y.data() + start * ell);
    }
    for (double x : y) if (!std::isfinite(x)) throw std::overflow_error("non-finite randomized ID sketch");
    std::vector<int64_t> jpvt(n);
    std::vector<double> tau(std::min(ell, n));
    geqp3(ell, n, y.data(), ell, jpvt.data(), tau.data());
    const int64_t steps = std::min(ell, n);
    std::vector<double> energy(steps), diag(steps), tails(steps + 1, 0.0);
    for (int64_t i = 0; i < steps; ++i) {
        double sum = 0;
        for (int64_t g = i; g < n; ++g) { const double r = y[g * ell + i]; sum += r * r; }
        energy[i] = sum;
        diag[i] = std::abs(y[i * ell + i]);
    }
    double diag_max = 0;
    for (int64_t i = steps; i-- > 0;) {
        tails[i] = std::hypot(tails[i + 1], std::sqrt(energy[i]));
        diag_max = std::max(diag_max, diag[i]);
    }
    if (!std::isfinite(tails[0]) || !std::isfinite(diag_max))
        throw std::overflow_error("non-finite randomized ID QR residual");
    const double norm = tails[0];
    const double floor = EPS * (double)std::max(ell, n) * diag_max;
    int64_t numerical_rank = steps;
    for (int64_t i = 0; i < steps; ++i) if (diag[i] <= floor) { numerical_rank = i; break; }
    const int64_t cap = std::min({max_rank ? max_rank : ell - oversampling, ell - oversampling, steps,
                                  physical_rank, numerical_rank});
    if (max_rank && numerical_rank < max_rank)
        throw std::invalid_argument("requested anchor_rank exceeds the numerical rank of the randomized sketch");
    int64_t rank = max_rank ? max_rank : cap;
    if (!max_rank)
        for (int64_t i = 0; i <= cap; ++i) if (tails[i] <= tolerance * norm) { rank = i; break; }
    for (int64_t i = 0; i < rank; ++i) res.selected.push_back(kept.original[jpvt[i] - 1]);
    res.norm = norm;
    res.numerical_rank = numerical_rank;
    res.cap = cap;
    return res;
}

template <class F>
int guarded(F&& f)
{
    try { f(); return 0; }
    catch (const std::invalid_argument& e) { g_sel_error = e.what(); return 1; }
    catch (const std::overflow_error& e) { g_sel_error = e.what(); return 2; }
    catch (const std::exception& e) { g_sel_error = e.what(); return 3; }
}

}  // namespace

extern "C" {

const char* oracle_select_last_error() { return g_sel_error.c_str(); }

int oracle_geqp3(int64_t m, int64_t n, double* a, int64_t lda, int64_t* jpvt, double* tau)
{
    return guarded([&] { geqp3(m, n, a, lda, jpvt, tau); });
}

int oracle_geqp3_steps(int64_t m, int64_t n, double* a, int64_t lda, int64_t* jpvt, double* tau, int64_t kmax)
{
    return guarded([&] { geqp3(m, n, a, lda, jpvt, tau, kmax); });
}

int oracle_omega(int64_t d, int64_t ell, uint64_t seed, double* u, double* v)
{
    return guarded([&] { omega(d, ell, seed, u, v); });
}

int oracle_sketch(const double* phi, int64_t count, int64_t d, const double* u, const double* v,
                  int64_t ell, const double* sw, double* y)
{
    return guarded([&] { sketch(phi, count, d, u, v, ell, sw, y); });
}

// algorithm: 0 dense QRCP, 1 online QRCP, 2 randomized quadrature KR sketch,
// 4 / 5 AO-density KR sketches, raw / normalized (LocalAnchorSelectionAlgorithm
// values, enums.hpp:23-30); importance[n] is read by 4 and 5 only.
int oracle_select(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                  const double* env, int64_t nenv, int cart,
                  const double* points, const double* weights, int64_t n,
                  const int32_t* axes, int naxes, const double* mo_coeff, int64_t nocc,
                  const double* axis_weights, double tolerance, int64_t max_rank, int algorithm,
                  int64_t batch_size, int64_t sketch_size, int64_t oversampling, uint64_t seed,
                  const double* importance, double screening,
                  int64_t* out, int64_t capacity, int64_t* count, int64_t* stats)
{
    return guarded([&] {
        Tables t{atm, natm, bas, nbas, env, nenv, cart, axes, naxes, mo_coeff, nocc};
        if (axis_weights)
            for (int a = 0; a < naxes; ++a)
                if (!std::isfinite(axis_weights[a]) || axis_weights[a] <= 0)
                    throw std::invalid_argument("axis weights must be finite and positive");
        std::vector<double> scales;
        if (naxes > 1) {
            scales.resize(naxes);
            if (oracle_axis_scales(atm, natm, bas, nbas, env, nenv, cart, points, weights, n, axes, naxes,
                                   mo_coeff, nocc, axis_weights, scales.data()))
                throw std::invalid_argument(oracle_last_error());
        }
        std::vector<int64_t> sel;
        if (algorithm == 0) sel = select_dense(t, points, weights, n, tolerance, max_rank, scales);
        else if (algorithm == 1) sel = select_online(t, points, weights, n, tolerance, max_rank, batch_size, scales, stats);
        else if (algorithm == 2 || algorithm == 4 || algorithm == 5)
            sel = select_rand(t, points, weights, n, tolerance, max_rank, batch_size, sketch_size, oversampling, seed,
                              scales, algorithm == 2 ? 0 : algorithm - 3, importance, screening).selected;
        else throw std::invalid_argument("the oracle restates qrcp_dense, qrcp_online and the randomized sketches");
        if ((int64_t)sel.size() > capacity) throw std::invalid_argument("output capacity too small");
        std::copy(sel.begin(), sel.end(), out);
        *count = (int64_t)sel.size();
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CVT anchors: cpp/src/anchor_select_cvt.cpp:17-204 (host order of operations:
// mass-weighted k-means++ seeding with long-double prefix sampling, Lloyd
// iterations with centroid sums in point order, empty clusters reseeded at
// the largest residual, snapping to distinct sites, sorted indices).
// ---------------------------------------------------------------------------
namespace {

std::vector<int64_t> cvt_run(const double* points, const std::vector<int64_t>& active,
                             const std::vector<double>& mass, int64_t rank, uint64_t seed, int64_t max_iter,
                             double tol)
{
    const int64_t n = (int64_t)active.size();
    auto P = [&](int64_t i, int d) { return points[3 * active[i] + d]; };
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
    std::vector<std::array<double, 3>> c(rank);
    const int64_t first = sample(mass);
    for (int d = 0; d < 3; ++d) c[0][d] = P(first, d);
    std::vector<double> nearest(n), scores(n);
    for (int64_t k = 1; k < rank; ++k) {
        for (int64_t i = 0; i < n; ++i) {
            double d2 = 0;
            for (int d = 0; d < 3; ++d) {
                const double delta = P(i, d) - c[k - 1][d];
                d2 += delta * delta;
            }
            nearest[i] = k == 1 ? d2 : (d2 < nearest[i] ? d2 : nearest[i]);
        }
        long double total = 0;
        for (int64_t i = 0; i < n; ++i) {
            scores[i] = mass[i] * nearest[i];
            if (!std::isfinite(scores[i])) throw std::invalid_argument("CVT weighted distance overflow");
            total += scores[i];
        }
        const int64_t pick = sample(total > 0 ? scores : mass);
        for (int d = 0; d < 3; ++d) c[k][d] = P(pick, d);
    }
    std::array<double, 3> lo, hi;
    for (int d = 0; d < 3; ++d) lo[d] = hi[d] = P(0, d);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], P(i, d));
            hi[d] = std::max(hi[d], P(i, d));
        }
    double scale2 = 0;
    for (int d = 0; d < 3; ++d) scale2 += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    const double scale = std::max(std::sqrt(scale2), 1.0);
    std::vector<double> sums(4 * rank), residual(n);
    for (int64_t iteration = 0; iteration < max_iter; ++iteration) {
        std::fill(sums.begin(), sums.end(), 0.0);
        for (int64_t i = 0; i < n; ++i) {
            double best = 1.7976931348623157e308;
            int64_t label = 0;
            for (int64_t k = 0; k < rank; ++k) {
                double d2 = 0;
                for (int d = 0; d < 3; ++d) {
                    const double delta = P(i, d) - c[k][d];
                    d2 += delta * delta;
                }
                if (d2 < best) { best = d2; label = k; }
            }
            const double m = mass[i];
            sums[4 * label] += m;
            for (int d = 0; d < 3; ++d) sums[4 * label + 1 + d] += m * P(i, d);
            residual[i] = m * best;
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
                if (!reseeded[i] && residual[i] > best) { best = residual[i]; site = i; }
            reseeded[site] = 1;
            for (int d = 0; d < 3; ++d) next[k][d] = P(site, d);
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
    std::vector<unsigned char> used(n, 0);
    std::vector<int64_t> indices;
    for (const auto& center : c) {
        int64_t site = n;
        double best = std::numeric_limits<double>::infinity();
        for (int64_t i = 0; i < n; ++i) {
            if (used[i]) continue;
            double d2 = 0;
            for (int d = 0; d < 3; ++d) {
                const double delta = P(i, d) - center[d];
                d2 += delta * delta;
            }
            if (d2 < best || (d2 == best && (site == n || active[i] < active[site]))) { best = d2; site = i; }
        }
        used[site] = 1;
        indices.push_back(active[site]);
    }
    std::sort(indices.begin(), indices.end());
    return indices;
}

}  // namespace

extern "C" int oracle_select_cvt(const double* points, const double* weights, const double* importance, int64_t npts,
                                 int64_t rank, uint64_t seed, int64_t max_iter, double tol, int64_t* out,
                                 int64_t* count)
{
    return guarded([&] {
        if (rank <= 0 || max_iter <= 0 || !std::isfinite(tol) || tol < 0)
            throw std::invalid_argument("CVT requires positive rank and max_iter, finite nonnegative tol");
        std::vector<int64_t> active;
        std::vector<double> mass;
        for (int64_t i = 0; i < npts; ++i) {
            for (int d = 0; d < 3; ++d)
                if (!std::isfinite(points[3 * i + d])) throw std::invalid_argument("CVT coordinates must be finite");
            if (!std::isfinite(weights[i]) || !std::isfinite(importance[i]) || importance[i] < 0)
                throw std::invalid_argument("CVT weights must be finite and nonnegative");
            const double m = weights[i] == 0 ? 0.0 : importance[i];
            if (m > 0) { active.push_back(i); mass.push_back(m); }
        }
        if (rank > (int64_t)active.size())
            throw std::invalid_argument("CVT requires at least rank positive-importance sites");
        const auto sel = cvt_run(points, active, mass, rank, seed, max_iter, tol);
        std::copy(sel.begin(), sel.end(), out);
        *count = (int64_t)sel.size();
    });
}

// libstdc++ std::discrete_distribution draws on std::mt19937_64, the sampler of
// the reference's RBRP blocks (cpp/src/anchor_select_rbrp.cpp:246-256), for
// the NumPy restatement in oracle/__init__.py.
extern "C" void* oracle_rng_create(uint64_t seed) { return new std::mt19937_64(seed); }
extern "C" void oracle_rng_destroy(void* rng) { delete static_cast<std::mt19937_64*>(rng); }
extern "C" int64_t oracle_rng_discrete(void* rng, const double* weights, int64_t n)
{
    std::discrete_distribution<size_t> dist(weights, weights + n);
    return (int64_t)dist(*static_cast<std::mt19937_64*>(rng));
}
// std::uniform_real_distribution<double>(0, total) on the same engine: the
// draw of the reference's device RBRP sampler
// (cpp/src/anchor_select_rbrp_device.cpp:129-160).
extern "C" double oracle_rng_uniform(void* rng, double total)
{
    std::uniform_real_distribution<double> dist(0.0, total);
    return dist(*static_cast<std::mt19937_64*>(rng));
}
