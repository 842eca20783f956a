// qrcp_online on the GPU: bounded-memory multipass Householder QRCP over the
// weighted pair columns a_g = sqrt|w_g| vec(f_g f_g^T).
//
// The decision logic is the reference's, step for step
// (cpp/src/anchor_select_online_qr.cpp:175-375): initial norms, thresholds
// floor = 32 eps ||norms||, refresh sqrt(eps), gamma = m eps / (1 - m eps);
// argmax with lowest index on ties; direct rescans before termination;
// exact candidate verification with reselection; host Householder reflector
// (make_reflector, :111-123); q_k = H_0..H_k e_k; norm downdate with the
// same refresh triggers.  What changes is where the work runs: features stay
// resident in HBM (no AO re-evaluation per pass), every pass is one GEMM
// F M plus row dots (quadratic forms instead of d^2-length columns), and
// explicit columns (checks, refreshes, rescans) are transformed by one
// workgroup each.
#include "online.hpp"

#include "online_kernels.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>
#include <limits>

namespace mpisdf {

std::vector<int64_t> select_online(Context& ctx, const SelectArgs& a, const AxisSpec& spec,
                                   const std::vector<double>& scales, double* diag, hipStream_t s)
{
    const auto started = std::chrono::steady_clock::now();
    if (!std::isfinite(a.tolerance) || a.tolerance < 0 || a.batch_size == 0)
        throw std::invalid_argument("invalid online QR weights, tolerance or batch size");
    const int64_t width = spec.width;
    if (width > (1 << 15)) throw std::overflow_error("AO pair dimension is too large");
    std::vector<int64_t> original;
    std::vector<double> kept_xyz, hsw;
    for (int64_t i = 0; i < a.n; ++i) {
        if (!std::isfinite(a.weights[i])) throw std::invalid_argument("grid weights must be finite");
        if (a.weights[i] == 0) continue;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(a.points[3 * i + c])) throw std::invalid_argument("grid coordinates must be finite");
        original.push_back(i);
        kept_xyz.insert(kept_xyz.end(), a.points + 3 * i, a.points + 3 * i + 3);
        hsw.push_back(std::sqrt(std::abs(a.weights[i])));
    }
    if (original.empty()) return {};
    const int64_t m = width * width, n = (int64_t)original.size();
    const int64_t cap = std::min({m, n, a.max_rank ? a.max_rank : std::min(m, n)});
    constexpr double eps = std::numeric_limits<double>::epsilon();

    // Resident features, |features| and weights.
    DeviceBuffer<double> xyz, feat((size_t)(n * width)), absfeat((size_t)(n * width)), sw;
    xyz.upload(kept_xyz.data(), kept_xyz.size(), s);
    sw.upload(hsw.data(), hsw.size(), s);
    {
        AxisWorkspace ws;
        const int64_t batch = std::max<int64_t>(1024, std::min<int64_t>(n, (int64_t(1) << 30) / (8 * (width + 4 * ctx.nao))));
        for (int64_t start = 0; start < n; start += batch) {
            const int64_t count = std::min(batch, n - start);
            evaluate_axis_device(ctx, spec, xyz.ptr + 3 * start, count, scales.empty() ? nullptr : scales.data(),
                                 a.mo_coeff, a.nocc, feat.ptr + start * width, ws, s);
        }
    }
    online_abs(feat.ptr, n * width, absfeat.ptr, s);

    // Reflector storage grows with the rank (capacity doubling as the reference).
    int64_t capacity = std::min<int64_t>(cap, 8);
    DeviceBuffer<double> refl((size_t)(std::max<int64_t>(capacity, 1) * m)), taus(std::max<int64_t>(capacity, 1));
    std::vector<double> tau_host;
    const int64_t scratch_cols = std::max<int64_t>(1, std::min<int64_t>(n, std::max<int64_t>(64, (int64_t(1) << 28) / m)));
    DeviceBuffer<double> scratch((size_t)(scratch_cols * m));
    DeviceBuffer<int64_t> d_ids(n);
    DeviceBuffer<double> d_norms(n), d_col(m), qv(m), absq(m), tmat((size_t)(n * width)), row(n), absrow(n);
    std::vector<int64_t> all_ids(n);
    for (int64_t i = 0; i < n; ++i) all_ids[i] = i;

    std::vector<double> estimated(n), reference(n);
    std::vector<unsigned char> selected(n, 0);
    std::vector<int64_t> result, result_local;
    int64_t full_scans = 0, candidate_checks = 0, reselections = 0, norm_updates = 0, refreshed = 0;

    // Tail norms of explicit columns (local ids) after `rank` reflectors, from row `first`.
    auto explicit_norms = [&](const std::vector<int64_t>& ids, int64_t rank, int64_t first, double* column_out) {
        std::vector<double> out(ids.size());
        for (size_t off = 0; off < ids.size(); off += (size_t)n) {
            const int64_t cnt = (int64_t)std::min(ids.size() - off, (size_t)n);
            MPISDF_HIP(hipMemcpyAsync(d_ids.ptr, ids.data() + off, cnt * sizeof(int64_t), hipMemcpyHostToDevice, s));
            online_transform_norms(feat.ptr, sw.ptr, width, d_ids.ptr, cnt, refl.ptr, taus.ptr, rank, first,
                                   scratch.ptr, scratch_cols, d_norms.ptr, column_out ? d_col.ptr : nullptr, s);
            MPISDF_HIP(hipMemcpyAsync(out.data() + off, d_norms.ptr, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        MPISDF_HIP(hipStreamSynchronize(s));
        for (double v : out)
            if (!std::isfinite(v)) throw std::overflow_error("non-finite online AO pair column");
        return out;
    };
    auto scan = [&](bool direct) {
        if (direct) ++full_scans;
        const int64_t rank = (int64_t)result.size();
        std::vector<double> norms;
        if (!direct && rank == 0) {
            // ||a_g|| = sqrt|w_g| ||f_g||^2 for every column (row pass with q unused).
            norms = explicit_norms(all_ids, 0, 0, nullptr);
        } else {
            norms = explicit_norms(all_ids, direct ? rank : 0, rank, nullptr);
        }
        for (int64_t g = 0; g < n; ++g) {
            estimated[g] = selected[g] ? 0 : norms[g];
            reference[g] = estimated[g];
        }
    };
    scan(false);
    const double first = *std::max_element(estimated.begin(), estimated.end());
    double frob = 0;
    for (double x : estimated) frob = std::hypot(frob, x);
    const double floor = 32 * eps * frob;
    const double refresh_threshold = std::sqrt(eps);
    const double gamma = ((double)m * eps < 0.5) ? (double)m * eps / (1 - (double)m * eps) : 1.0;
    auto winner = [&]() {
        int64_t p = 0;
        for (int64_t g = 1; g < n; ++g)
            if (estimated[g] > estimated[p]) p = g;
        return p;
    };
    std::vector<double> column(m);
    while (true) {
        int64_t p = winner();
        if (estimated[p] <= a.tolerance * first || (int64_t)result.size() == cap || estimated[p] <= floor) {
            scan(true);  // no optimistic estimate may terminate the algorithm
            p = winner();
            if (estimated[p] <= a.tolerance * first || (int64_t)result.size() == cap || estimated[p] <= floor) break;
        }
        for (;;) {
            p = winner();
            const double norm = explicit_norms({p}, (int64_t)result.size(), (int64_t)result.size(), column.data())[0];
            ++candidate_checks;
            estimated[p] = reference[p] = norm;
            if (p == winner()) {
                MPISDF_HIP(hipMemcpyAsync(column.data(), d_col.ptr, m * sizeof(double), hipMemcpyDeviceToHost, s));
                MPISDF_HIP(hipStreamSynchronize(s));
                break;
            }
            ++reselections;
        }
        if (estimated[p] <= floor || estimated[p] <= a.tolerance * first) {
            scan(true);
            if (estimated[winner()] <= floor || estimated[winner()] <= a.tolerance * first) break;
            continue;
        }
        const int64_t k = (int64_t)result.size();
        double tau = 0;
        {  // make_reflector (online_qr.cpp:111-123)
            double tail = 0;
            for (int64_t i = k + 1; i < m; ++i) tail = std::hypot(tail, column[i]);
            const double alpha = column[k];
            if (tail == 0) {
                tau = 0;
                column[k] = 1;
            } else {
                const double beta = -std::copysign(std::hypot(alpha, tail), alpha);
                tau = (beta - alpha) / beta;
                const double divisor = alpha - beta;
                column[k] = 1;
                for (int64_t i = k + 1; i < m; ++i) column[i] /= divisor;
            }
            for (int64_t i = 0; i < k; ++i) column[i] = 0;
        }
        if (k == capacity) {
            const int64_t next = std::min(cap, std::max<int64_t>(1, 2 * capacity));
            DeviceBuffer<double> nr((size_t)(next * m)), nt(next);
            MPISDF_HIP(hipMemcpyAsync(nr.ptr, refl.ptr, k * m * sizeof(double), hipMemcpyDeviceToDevice, s));
            MPISDF_HIP(hipMemcpyAsync(nt.ptr, taus.ptr, k * sizeof(double), hipMemcpyDeviceToDevice, s));
            MPISDF_HIP(hipStreamSynchronize(s));
            refl = std::move(nr);
            taus = std::move(nt);
            capacity = next;
        }
        MPISDF_HIP(hipMemcpyAsync(refl.ptr + k * m, column.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
        MPISDF_HIP(hipMemcpyAsync(taus.ptr + k, &tau, sizeof(double), hipMemcpyHostToDevice, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        tau_host.push_back(tau);
        result.push_back(original[p]);
        result_local.push_back(p);
        selected[p] = 1;
        estimated[p] = reference[p] = 0;

        online_make_q(refl.ptr, taus.ptr, m, k, qv.ptr, s);
        online_row_pass(feat.ptr, absfeat.ptr, sw.ptr, n, width, qv.ptr, absq.ptr, tmat.ptr, row.ptr, absrow.ptr, s);
        std::vector<double> hrow(n), habs(n);
        MPISDF_HIP(hipMemcpyAsync(hrow.data(), row.ptr, n * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipMemcpyAsync(habs.data(), absrow.ptr, n * sizeof(double), hipMemcpyDeviceToHost, s));
        MPISDF_HIP(hipStreamSynchronize(s));
        std::vector<int64_t> refresh;
        for (int64_t id = 0; id < n; ++id) {
            if (selected[id]) continue;
            ++norm_updates;
            const double old = estimated[id], r = std::abs(hrow[id]);
            if (!std::isfinite(r) || !std::isfinite(habs[id]))
                throw std::overflow_error("non-finite online QR coefficient");
            const double ratio = old > 0 ? r / old : 0;
            const double temp = std::max(0.0, (1 + ratio) * (1 - ratio));
            const double next = old * std::sqrt(temp);
            const double ref_ratio = reference[id] > 0 ? old / reference[id] : 0;
            const double err = gamma * habs[id];
            estimated[id] = next;
            if (old == 0 || ratio > 1 || temp * ref_ratio * ref_ratio <= refresh_threshold ||
                (2 * r * err + err * err) > refresh_threshold * next * next)
                refresh.push_back(id);
        }
        if (!refresh.empty()) {
            refreshed += (int64_t)refresh.size();
            const auto nr = explicit_norms(refresh, k + 1, k + 1, nullptr);
            for (size_t j = 0; j < refresh.size(); ++j) estimated[refresh[j]] = reference[refresh[j]] = nr[j];
        }
    }
    if (a.debug) {
        const double residual = *std::max_element(estimated.begin(), estimated.end());
        std::cerr << "[online qrcp] rank=" << result.size() << " rows=" << m << " columns=" << n
                  << " batch_size=" << a.batch_size << " max_initial_norm=" << first << " max_residual=" << residual
                  << " relative_residual=" << (first > 0 ? residual / first : 0) << " full_scans=" << full_scans
                  << " candidate_checks=" << candidate_checks << " candidate_reselections=" << reselections
                  << " norm_updates=" << norm_updates << " refreshed_columns=" << refreshed << " elapsed="
                  << std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count() << " s\n";
    }
    if (diag) {
        diag[kDiagRows] = (double)m;
        diag[kDiagColumns] = (double)n;
        diag[kDiagCap] = (double)cap;
        diag[kDiagRank] = (double)result.size();
        diag[kDiagBlocks] = (double)full_scans;
        diag[kDiagFolds] = (double)refreshed;
    }
    return result;
}

}  // namespace mpisdf
