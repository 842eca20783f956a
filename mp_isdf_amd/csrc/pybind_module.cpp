// pybind11 module ``mp_isdf``: the drop-in Python surface of the reference
// extension (cpp/src/ao_eval_pybind.cpp:70-120, cpp/src/anchor_select_pybind.cpp:93-233).
//
// Host C++ only: every computation goes through the C ABI of
// libmp_isdf_amd.so (include/mp_isdf_amd.h), which runs on the GPU.  This
// build has no host execution space: ``space="host"`` raises ValueError, the
// same way the reference raises for ``space="device"`` in a build without
// its accelerator backend (cpp/src/ao_eval.cpp:293-295).
#include "../../include/mp_isdf_amd.h"

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace py = pybind11;

namespace {

using IntArray = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;
using FloatArray = py::array_t<double, py::array::c_style | py::array::forcecast>;

enum class AxisType { AtomicOrbital, AtomicOrbitalGradX, AtomicOrbitalGradY, AtomicOrbitalGradZ, OccupiedOrbital };
enum class LocalAlgorithm {
    QRColumnPivoting,
    OnlineQRColumnPivoting,
    RandomizedQuadratureKhatriRao,
    CentroidalVoronoi,
    RandomizedAODensityKhatriRao,
    RandomizedAODensityNormalizedKhatriRao
};
enum class DistributedAlgorithm { SVD, RandomizedQRColumnPivoting };

void check(int status)
{
    if (status == 0) return;
    const std::string msg = mpisdf_last_error();
    if (status == 1) throw py::value_error(msg);
    if (status == 2) { PyErr_SetString(PyExc_OverflowError, msg.c_str()); throw py::error_already_set(); }
    throw std::runtime_error(msg);
}

// Execution spaces of the reference (cpp/src/evaluation_space_pybind.hpp:8-14).
void parse_space(const std::string& space)
{
    if (space == "auto" || space == "default" || space == "device") return;
    if (space == "host")
        throw py::value_error("this mp_isdf build runs on the GPU only; space='host' is not available "
                              "(use 'auto' or 'device')");
    throw py::value_error("space must be 'auto', 'host', or 'device'");
}

bool device_available()
{
    int count = 0;
    return mpisdf_device_count(&count) == 0 && count > 0;
}

class Context {
public:
    Context(const IntArray& atm, const IntArray& bas, const FloatArray& env, bool cart)
    {
        if (atm.ndim() != 2 || atm.shape(1) != 6 || bas.ndim() != 2 || bas.shape(1) != 8 || env.ndim() != 1)
            throw py::value_error("expected atm[natm,6], bas[nbas,8], env[nenv]");
        mpisdf_context* raw = nullptr;
        check(mpisdf_context_create(atm.data(), atm.shape(0), bas.data(), bas.shape(0), env.data(),
                                    env.shape(0), cart ? 1 : 0, &raw));
        handle_.reset(raw, [](mpisdf_context* c) { mpisdf_context_destroy(c); });
        int64_t n = 0;
        check(mpisdf_context_nao(raw, &n));
        nao_ = n;
        natm_ = atm.shape(0);
        cart_ = cart;
    }
    int64_t nao() const { return nao_; }
    int64_t natm() const { return natm_; }
    bool cart() const { return cart_; }
    const mpisdf_context* get() const { return handle_.get(); }
    mpisdf_context* get_mut() { return handle_.get(); }

    py::array_t<double> evaluate(const FloatArray& coords, bool gradient, const std::string& space) const
    {
        parse_space(space);
        if (coords.ndim() != 2 || coords.shape(1) != 3)
            throw py::value_error("expected coords[ngrid,3] in Bohr");
        const py::ssize_t n = coords.shape(0);
        std::vector<py::ssize_t> shape = gradient ? std::vector<py::ssize_t>{3, n, (py::ssize_t)nao_}
                                                  : std::vector<py::ssize_t>{n, (py::ssize_t)nao_};
        py::array_t<double> out(shape);
        if (n == 0 || nao_ == 0) return out;
        const double* src = coords.data();
        double* dst = out.mutable_data();
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_eval_ao(get(), src, n, gradient ? 1 : 0, dst);
        }
        check(status);
        return out;
    }

private:
    std::shared_ptr<mpisdf_context> handle_;
    int64_t nao_ = 0, natm_ = 0;
    bool cart_ = false;
};

std::vector<int32_t> read_axes(py::object axes)
{
    std::vector<AxisType> list;
    if (py::isinstance<py::list>(axes) || py::isinstance<py::tuple>(axes)) list = axes.cast<std::vector<AxisType>>();
    else list = {axes.cast<AxisType>()};
    std::vector<int32_t> codes;
    for (auto a : list) codes.push_back(static_cast<int32_t>(a));
    return codes;
}

std::vector<double> read_coefficients(const Context& ctx, py::object mo_coeff)
{
    if (mo_coeff.is_none()) return {};
    const auto c = mo_coeff.cast<FloatArray>();
    const auto nao = static_cast<py::ssize_t>(ctx.nao());
    if (c.ndim() != 2 || c.shape(0) != nao || c.shape(1) != nao)
        throw py::value_error("mo_coeff must have shape (nao,nao)");
    return {c.data(), c.data() + c.size()};
}

const double* read_points(const FloatArray& points)
{
    if (points.ndim() != 2 || points.shape(1) != 3) throw py::value_error("expected points[ngrid,3]");
    return points.data();
}

class Selector {
public:
    explicit Selector(const Context& ctx) : ctx_(ctx) {}
    const Context& context() const { return ctx_; }

    py::array_t<double> evaluate_axis(const FloatArray& points, py::object axes, py::object mo_coeff,
                                      int64_t nocc, const std::string& space) const
    {
        parse_space(space);
        const double* xyz = read_points(points);
        const auto codes = read_axes(axes);
        const auto coeff = read_coefficients(ctx_, mo_coeff);
        int64_t width = 0;
        check(mpisdf_evaluate_axis(ctx_.get(), xyz, 0, codes.data(), (int32_t)codes.size(),
                                   coeff.empty() ? nullptr : coeff.data(), nocc, nullptr, &width));
        py::array_t<double> out({points.shape(0), (py::ssize_t)width});
        if (points.shape(0) == 0) return out;
        double* dst = out.mutable_data();
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_evaluate_axis(ctx_.get(), xyz, points.shape(0), codes.data(), (int32_t)codes.size(),
                                          coeff.empty() ? nullptr : coeff.data(), nocc, dst, &width);
        }
        check(status);
        return out;
    }

    std::vector<int64_t> select(const FloatArray& points, const FloatArray& weights, py::object axes,
                                double tolerance, int64_t max_rank, LocalAlgorithm algorithm,
                                py::object mo_coeff, int64_t nocc, const std::string& space,
                                int64_t batch_size, bool debug, int64_t sketch_size, int64_t oversampling,
                                uint64_t seed, bool audit, py::object cvt_weights, int64_t cvt_max_iter,
                                double cvt_tol, const std::vector<double>& axis_weights,
                                py::object point_importance, double ao_screening_threshold,
                                const std::string& precision, double* diagnostics = nullptr) const
    {
        parse_space(space);
        const double* xyz = read_points(points);
        if (weights.ndim() != 1 || weights.shape(0) != points.shape(0))
            throw py::value_error("expected weights[ngrid]");
        const auto codes = read_axes(axes);
        const auto coeff = read_coefficients(ctx_, mo_coeff);
        if (algorithm == LocalAlgorithm::CentroidalVoronoi) {
            // anchor_select.cpp:167-169: CVT reads the grid, the quadrature
            // weights and cvt_weights only.
            if (cvt_weights.is_none()) throw py::value_error("CVT requires cvt_weights (ngrid,)");
            const auto cw = cvt_weights.cast<FloatArray>();
            if (cw.ndim() != 1 || cw.shape(0) != points.shape(0))
                throw py::value_error("CVT points, quadrature weights and CVT weights must have equal lengths");
            if (max_rank < 0 || cvt_max_iter < 0) throw py::value_error("sizes must be non-negative");
            std::vector<int64_t> out(std::max<int64_t>(max_rank, 1));
            int64_t count = 0;
            int status;
            {
                py::gil_scoped_release release;
                status = mpisdf_select_cvt(xyz, weights.data(), cw.data(), points.shape(0), max_rank, seed,
                                           cvt_max_iter, cvt_tol, out.data(), (int64_t)out.size(), &count);
            }
            check(status);
            out.resize(count);
            return out;
        }
        // anchor_select_pybind.cpp:186-191: the importance, when given, must
        // match the grid; only the density sketches use it.
        std::vector<double> importance;
        if (!point_importance.is_none()) {
            const auto input = point_importance.cast<FloatArray>();
            if (input.ndim() != 1 || input.shape(0) != points.shape(0))
                throw py::value_error("point_importance must have shape (ngrid,)");
            importance.assign(input.data(), input.data() + input.size());
        }
        if (batch_size < 0 || sketch_size < 0 || oversampling < 0 || max_rank < 0 || nocc < 0)
            throw py::value_error("sizes must be non-negative");
        int prec = 0;
        if (precision == "fp64") prec = 0;
        else if (precision == "fp32") prec = 1;
        else if (precision == "bf16") prec = 2;
        else if (precision == "fp8") prec = 3;
        else if (precision == "bf16x3") prec = 4;
        else if (precision == "fp64i8") prec = 5;
        else throw py::value_error("precision must be 'fp64', 'fp64i8', 'fp32', 'bf16x3', 'bf16' or 'fp8'");
        std::vector<int64_t> out(std::max<py::ssize_t>(points.shape(0), 1));
        int64_t count = 0;
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_select_importance(
                ctx_.get(), xyz, weights.data(), points.shape(0), codes.data(), (int32_t)codes.size(),
                coeff.empty() ? nullptr : coeff.data(), nocc, axis_weights.empty() ? nullptr : axis_weights.data(),
                (int64_t)axis_weights.size(), tolerance, max_rank, static_cast<int32_t>(algorithm), batch_size,
                (debug ? 1 : 0) | (audit ? 2 : 0), sketch_size, oversampling, seed, prec, importance.empty() ? nullptr : importance.data(),
                ao_screening_threshold, out.data(), (int64_t)out.size(), &count, diagnostics);
        }
        check(status);
        out.resize(count);
        return out;
    }

    py::tuple select_with_diagnostics(const FloatArray& points, const FloatArray& weights, py::object axes,
                                      double tolerance, int64_t max_rank, LocalAlgorithm algorithm,
                                      py::object mo_coeff, int64_t nocc, const std::string& space, int64_t batch_size,
                                      bool debug, int64_t sketch_size, int64_t oversampling, uint64_t seed,
                                      const std::vector<double>& axis_weights, const std::string& precision,
                                      bool audit, py::object point_importance, double ao_screening_threshold) const
    {
        std::vector<double> diag(28, 0.0);
        auto idx = select(points, weights, axes, tolerance, max_rank, algorithm, mo_coeff, nocc, space, batch_size,
                          debug, sketch_size, oversampling, seed, audit, py::none(), 100, 1e-5, axis_weights,
                          point_importance, ao_screening_threshold, precision, diag.data());
        static const char* names[] = {"rows", "columns", "numerical_rank", "rank_cap", "rank", "sketch_norm",
                                      "relative_tail", "blocks", "folds", "seconds_sketch", "seconds_qrcp",
                                      "seconds_total", "stalls", "refreshes", "fold_seconds", "fold_flops",
                                      "fold_launches", "sketch_kernel_seconds", "sketch_flops", "sketch_launches",
                                      "fold_kind", "sketch_int8_ops", "gs_seconds", "gs_flops"};
        py::dict d;
        for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); ++i) d[names[i]] = diag[i];
        // The audit's residuals (selected, dense QRCP) per objective, when it ran.
        if (diag[24] >= 0) d["audit"] = py::make_tuple(diag[24], diag[25]);
        if (diag[26] >= 0) d["audit_density"] = py::make_tuple(diag[26], diag[27]);
        return py::make_tuple(idx, d);
    }

private:
    const Context& ctx_;
};

}  // namespace

PYBIND11_MODULE(mp_isdf, m)
{
    m.doc() = "MI355X (gfx950) AO evaluation and ISDF anchor selection (drop-in for the reference mp_isdf)";
    const bool dev = device_available();
    m.attr("device_enabled") = dev;
    // Reference name of the same flag (cpp/src/ao_eval_pybind.cpp:73-77).
    m.attr("cuda_enabled") = dev;

    py::class_<Context>(m, "LibcintContext",
                        "Own a private copy of libcint tables and cache AO data on the GPU.")
        .def(py::init<const IntArray&, const IntArray&, const FloatArray&, bool>(), py::arg("atm"),
             py::arg("bas"), py::arg("env"), py::arg("cart") = false)
        .def_property_readonly("nao", &Context::nao)
        .def_property_readonly("cart", &Context::cart)
        .def("release_workspace", [](Context& c) { check(mpisdf_context_release_workspace(c.get_mut())); },
             "Give back the sketch storage the randomized selectors keep between calls.")
        .def("eval_ao", [](const Context& c, const FloatArray& coords, const std::string& space) {
            return c.evaluate(coords, false, space);
        }, py::arg("coords"), py::arg("space") = "auto", "Return AO values with shape (ngrid, nao).")
        .def("eval_ao_gradient", [](const Context& c, const FloatArray& coords, const std::string& space) {
            return c.evaluate(coords, true, space);
        }, py::arg("coords"), py::arg("space") = "auto",
           "Return x,y,z AO gradients with shape (3, ngrid, nao).")
        .def("quadrature_overlap", [](const Context& c, const FloatArray& coords, const FloatArray& weights,
                                      int64_t batch_size) {
            if (coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 || weights.shape(0) != coords.shape(0))
                throw py::value_error("expected coords[n,3] and weights[n]");
            py::array_t<double> s({(py::ssize_t)c.nao(), (py::ssize_t)c.nao()});
            int status;
            double* ps = s.mutable_data();
            {
                py::gil_scoped_release release;
                status = mpisdf_quadrature_overlap(c.get(), coords.data(), weights.data(), coords.shape(0), batch_size,
                                                   ps);
            }
            check(status);
            return s;
        }, py::arg("coords"), py::arg("weights"), py::arg("batch_size") = 65536,
           "Grid-quadrature AO overlap sum_g w_g phi_g phi_g^T (nao, nao), evaluated on the GPU.")
        .def("overlap", [](const Context& c) {
            py::array_t<double> s({(py::ssize_t)c.nao(), (py::ssize_t)c.nao()});
            check(mpisdf_overlap(c.get(), s.mutable_data()));
            return s;
        }, "Analytic AO overlap matrix (nao, nao) of the evaluated functions (the int1e_ovlp the reference's "
           "overlap check compares with).");

    m.def("eval_ao", [](const IntArray& atm, const IntArray& bas, const FloatArray& env,
                        const FloatArray& coords, const std::string& space) {
        parse_space(space);
        return Context(atm, bas, env, false).evaluate(coords, false, space);
    }, py::arg("atm"), py::arg("bas"), py::arg("env"), py::arg("coords"), py::arg("space") = "auto");
    m.def("eval_ao_gradient", [](const IntArray& atm, const IntArray& bas, const FloatArray& env,
                                 const FloatArray& coords, const std::string& space) {
        parse_space(space);
        return Context(atm, bas, env, false).evaluate(coords, true, space);
    }, py::arg("atm"), py::arg("bas"), py::arg("env"), py::arg("coords"), py::arg("space") = "auto");

    py::enum_<AxisType>(m, "AxisType")
        .value("AtomicOrbital", AxisType::AtomicOrbital)
        .value("AtomicOrbitalGradX", AxisType::AtomicOrbitalGradX)
        .value("AtomicOrbitalGradY", AxisType::AtomicOrbitalGradY)
        .value("AtomicOrbitalGradZ", AxisType::AtomicOrbitalGradZ)
        .value("OccupiedOrbital", AxisType::OccupiedOrbital);
    py::enum_<LocalAlgorithm>(m, "LocalAnchorSelectionAlgorithm")
        .value("QRColumnPivoting", LocalAlgorithm::QRColumnPivoting)
        .value("OnlineQRColumnPivoting", LocalAlgorithm::OnlineQRColumnPivoting)
        .value("RandomizedQuadratureKhatriRao", LocalAlgorithm::RandomizedQuadratureKhatriRao)
        .value("CentroidalVoronoi", LocalAlgorithm::CentroidalVoronoi)
        .value("RandomizedAODensityKhatriRao", LocalAlgorithm::RandomizedAODensityKhatriRao)
        .value("RandomizedAODensityNormalizedKhatriRao", LocalAlgorithm::RandomizedAODensityNormalizedKhatriRao);
    py::enum_<DistributedAlgorithm>(m, "DistributedAnchorSelectionAlgorithm")
        .value("SVD", DistributedAlgorithm::SVD)
        .value("RandomizedQRColumnPivoting", DistributedAlgorithm::RandomizedQRColumnPivoting);

    m.def("lebedev", [](int npoints) {
        py::array_t<double> xyz({(py::ssize_t)std::max(npoints, 0), (py::ssize_t)3}), w((py::ssize_t)std::max(npoints, 0));
        check(mpisdf_lebedev(npoints, xyz.mutable_data(), w.mutable_data()));
        return py::make_tuple(xyz, w);
    }, py::arg("npoints"),
       "The npoints-point Lebedev-Laikov rule: (unit vectors [n,3], weights [n] summing to 1).");
    m.def("becke_grid", [](const IntArray& charges, const FloatArray& coords, int level, bool device) {
        if (charges.ndim() != 1 || coords.ndim() != 2 || coords.shape(1) != 3 || coords.shape(0) != charges.shape(0))
            throw py::value_error("expected charges[natm] and coords[natm,3] (Bohr)");
        int64_t count = 0;
        check(mpisdf_becke_grid(charges.data(), coords.data(), charges.shape(0), level, device, nullptr, nullptr, 0,
                                &count));
        py::array_t<double> xyz({(py::ssize_t)count, py::ssize_t(3)}), w((py::ssize_t)count);
        int status;
        double* px = xyz.mutable_data();
        double* pw = w.mutable_data();
        {
            py::gil_scoped_release release;
            status = mpisdf_becke_grid(charges.data(), coords.data(), charges.shape(0), level, device, px, pw, count,
                                       &count);
        }
        check(status);
        return py::make_tuple(xyz, w);
    }, py::arg("charges"), py::arg("coords"), py::arg("level") = 3, py::arg("device") = true,
       "Becke-partitioned molecular grid (Bohr): returns (coords[n,3], weights[n]); the partition runs on the GPU "
       "when device is true and one is present (same weights as the host loop).  Not part of the reference "
       "extension, which receives PySCF's grid.");

    m.def("isdf_fit", [](const Context& ctx, const FloatArray& anchor_coords, const FloatArray& anchor_weights,
                         const FloatArray& coords, const FloatArray& weights, int64_t batch_size, bool interp_vecs,
                         bool weight_sums) {
        if (anchor_coords.ndim() != 2 || anchor_coords.shape(1) != 3 || anchor_weights.ndim() != 1 ||
            anchor_weights.shape(0) != anchor_coords.shape(0))
            throw py::value_error("expected anchor_coords[k,3] and anchor_weights[k]");
        if (coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 || weights.shape(0) != coords.shape(0))
            throw py::value_error("expected coords[n,3] and weights[n]");
        const py::ssize_t k = anchor_coords.shape(0), n = coords.shape(0);
        py::array_t<double> theta, tsum;
        double* pt = nullptr;
        double* ps = nullptr;
        if (interp_vecs) { theta = py::array_t<double>({k, n}); pt = theta.mutable_data(); }
        if (weight_sums) { tsum = py::array_t<double>(k); ps = tsum.mutable_data(); }
        std::vector<double> res(10, 0.0);
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_isdf_fit(ctx.get(), anchor_coords.data(), anchor_weights.data(), k, coords.data(),
                                     weights.data(), n, batch_size, pt, ps, res.data());
        }
        check(status);
        py::dict out;
        py::dict err;
        err["l2"] = res[0];
        err["relative"] = res[1];
        out["pair_space_error"] = err;
        py::dict diag;
        diag["condition_estimate"] = res[2];
        diag["gram_diag_max"] = res[3];
        diag["seconds_factor"] = res[4];
        diag["seconds_grid_pass"] = res[5];
        diag["retained_rank"] = (int64_t)res[8];
        diag["factorization"] = res[9] != 0 ? "pseudo-inverse" : "cholesky";
        out["fit_diagnostics"] = diag;
        out["interp_vecs"] = interp_vecs ? py::object(theta) : py::object(py::none());
        out["theta_weight_sums"] = weight_sums ? py::object(tsum) : py::object(py::none());
        return out;
    }, py::arg("context"), py::arg("anchor_coords"), py::arg("anchor_weights"), py::arg("coords"),
       py::arg("weights"), py::arg("batch_size") = 4096, py::arg("interp_vecs") = false,
       py::arg("weight_sums") = false,
       "GPU ISDF fit on anchors (normal equations of the anchor-pair least squares, FP64 Cholesky): "
       "pair-space error, optional interpolation vectors theta[k,n] and sum_g theta[:,g] w_g.");

    m.def("pair_space_error", [](const Context& ctx, const FloatArray& anchor_coords, const FloatArray& anchor_weights,
                                 const FloatArray& coords, const FloatArray& weights, const FloatArray& interp_vecs,
                                 int64_t batch_size) {
        const py::ssize_t k = anchor_coords.shape(0), n = coords.shape(0);
        if (anchor_coords.ndim() != 2 || anchor_coords.shape(1) != 3 || anchor_weights.ndim() != 1 ||
            anchor_weights.shape(0) != k || coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 ||
            weights.shape(0) != n || interp_vecs.ndim() != 2 || interp_vecs.shape(0) != k ||
            interp_vecs.shape(1) != n)
            throw py::value_error("expected anchor_coords[k,3], anchor_weights[k], coords[n,3], weights[n] and "
                                  "interp_vecs[k,n]");
        double res[2] = {0, 0};
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_pair_space_error(ctx.get(), anchor_coords.data(), anchor_weights.data(), k, coords.data(),
                                             weights.data(), n, interp_vecs.data(), batch_size, res);
        }
        check(status);
        py::dict out;
        out["l2"] = res[0];
        out["relative"] = res[1];
        return out;
    }, py::arg("context"), py::arg("anchor_coords"), py::arg("anchor_weights"), py::arg("coords"), py::arg("weights"),
       py::arg("interp_vecs"), py::arg("batch_size") = 4096,
       "Pair-space error {l2, relative} of given interpolation vectors (python/isdf.py:176-191), on the GPU.");

    m.def("isdf_fit_gradient", [](const Context& ctx, const FloatArray& anchor_coords,
                                  const FloatArray& anchor_weights, const FloatArray& coords,
                                  const FloatArray& weights, int64_t batch_size) {
        if (anchor_coords.ndim() != 2 || anchor_coords.shape(1) != 3 || anchor_weights.ndim() != 1 ||
            anchor_weights.shape(0) != anchor_coords.shape(0))
            throw py::value_error("expected anchor_coords[k,3] and anchor_weights[k]");
        if (coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 || weights.shape(0) != coords.shape(0))
            throw py::value_error("expected coords[n,3] and weights[n]");
        const py::ssize_t k = anchor_coords.shape(0), n = coords.shape(0);
        int64_t nao = 0;
        check(mpisdf_context_nao(ctx.get(), &nao));
        const py::ssize_t natm = (py::ssize_t)ctx.natm();
        py::array_t<double> gao({(py::ssize_t)3, (py::ssize_t)nao}), gat({natm, (py::ssize_t)3});
        double loss = 0;
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_isdf_fit_gradient(ctx.get(), anchor_coords.data(), anchor_weights.data(), k,
                                              coords.data(), weights.data(), n, batch_size, &loss,
                                              gao.mutable_data(), gat.mutable_data());
        }
        check(status);
        py::dict out;
        out["loss"] = loss;
        out["grad_ao"] = gao;
        out["grad_atoms"] = gat;
        return out;
    }, py::arg("context"), py::arg("anchor_coords"), py::arg("anchor_weights"), py::arg("coords"),
       py::arg("weights"), py::arg("batch_size") = 4096,
       "Pair-reconstruction loss of the FP64 fit (sum_g |w_g| ||pair_g - fit_g||^2) and its gradient with "
       "respect to the nuclear coordinates (grid and anchor points fixed), by the hand-written adjoint.");

    m.def("coulomb_energy_gradient", [](const Context& ctx, const FloatArray& anchor_coords, const FloatArray& coords,
                                        const FloatArray& weights, const FloatArray& dm, int64_t batch_size,
                                        bool gradient) {
        if (anchor_coords.ndim() != 2 || anchor_coords.shape(1) != 3)
            throw py::value_error("expected anchor_coords[k,3]");
        if (coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 || weights.shape(0) != coords.shape(0))
            throw py::value_error("expected coords[n,3] and weights[n]");
        int64_t nao = 0;
        check(mpisdf_context_nao(ctx.get(), &nao));
        if (dm.ndim() != 2 || dm.shape(0) != nao || dm.shape(1) != nao)
            throw py::value_error("dm must be a finite AO density matrix of shape (nao,nao)");
        const py::ssize_t k = anchor_coords.shape(0), n = coords.shape(0);
        const py::ssize_t natm = (py::ssize_t)ctx.natm();
        py::array_t<double> gao({(py::ssize_t)3, (py::ssize_t)nao}), gat({natm, (py::ssize_t)3});
        double energy = 0;
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_coulomb_energy_gradient(ctx.get(), anchor_coords.data(), k, coords.data(), weights.data(),
                                                    n, dm.data(), batch_size, &energy,
                                                    gradient ? gao.mutable_data() : nullptr,
                                                    gradient ? gat.mutable_data() : nullptr);
        }
        check(status);
        py::dict out;
        out["energy"] = energy;
        if (gradient) {
            out["grad_ao"] = gao;
            out["grad_atoms"] = gat;
        }
        return out;
    }, py::arg("context"), py::arg("anchor_coords"), py::arg("coords"), py::arg("weights"), py::arg("dm"),
       py::arg("batch_size") = 4096, py::arg("gradient") = true,
       "Coulomb energy E_J = tr(J dm)/2 of the ISDF fit on fixed anchors with the grid Coulomb metric, and "
       "(gradient=True) its gradient with respect to the nuclear coordinates, by the hand-written adjoint.");

    // Grid Coulomb metric (python/isdf.py:48-67, :76-92) on the GPU.
    auto metric = [](const FloatArray& left, const FloatArray& right, const FloatArray& coords,
                     const FloatArray& weights, bool same, bool symmetrize) {
        if (coords.ndim() != 2 || coords.shape(1) != 3 || weights.ndim() != 1 || weights.shape(0) != coords.shape(0))
            throw py::value_error("expected coords[n,3] and weights[n]");
        const py::ssize_t n = coords.shape(0);
        if (left.ndim() != 2 || left.shape(1) != n || right.ndim() != 2 || right.shape(1) != n)
            throw py::value_error("interpolation vectors must have shape (nvec, ngrid)");
        const py::ssize_t kl = left.shape(0), kr = right.shape(0);
        py::array_t<double> out({kl, kr});
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_grid_coulomb_metric(left.data(), kl, same ? left.data() : right.data(), kr,
                                                coords.data(), weights.data(), n, symmetrize ? 1 : 0,
                                                out.mutable_data());
        }
        check(status);
        return out;
    };
    m.def("grid_coulomb_metric", [metric](const FloatArray& interp_vecs, const FloatArray& coords,
                                          const FloatArray& weights) {
        return metric(interp_vecs, interp_vecs, coords, weights, true, true);
    }, py::arg("interp_vecs"), py::arg("coords"), py::arg("weights"),
       "Spherical-cell grid Coulomb metric of interp_vecs[nvec, ngrid] (not weight scaled), symmetrised.");
    m.def("grid_cross_metric", [metric](const FloatArray& left, const FloatArray& right, const FloatArray& coords,
                                        const FloatArray& weights) {
        return metric(left, right, coords, weights, false, false);
    }, py::arg("left"), py::arg("right"), py::arg("coords"), py::arg("weights"),
       "Grid Coulomb metric between two sets of interpolation vectors.");

    // Integral-direct Cholesky metric (cpp/src/coulomb_metric_pybind.cpp:16-92).
    auto cholesky_algorithm = [](const std::string& algorithm) -> int32_t {
        if (algorithm == "unblocked_greedy") return 0;
        if (algorithm == "unblocked_uniform") return 1;
        throw py::value_error("unknown Cholesky algorithm");
    };
    m.def("cholesky_eri", [cholesky_algorithm](const Context& ctx, double tolerance, int64_t max_rank,
                                               const std::string& algorithm, bool debug, uint64_t seed) {
        const int32_t alg = cholesky_algorithm(algorithm);
        mpisdf_eri_cholesky* h = nullptr;
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_cholesky_eri(ctx.get(), tolerance, max_rank, alg, seed, debug ? 1 : 0, &h);
        }
        check(status);
        std::unique_ptr<mpisdf_eri_cholesky, int (*)(mpisdf_eri_cholesky*)> owned(h, mpisdf_eri_cholesky_destroy);
        int64_t rank = 0, npair = 0;
        check(mpisdf_eri_cholesky_size(h, &rank, &npair));
        py::array_t<double> factor({(py::ssize_t)rank, (py::ssize_t)npair});
        py::array_t<double> residual((py::ssize_t)npair);
        py::array_t<py::ssize_t> pivots((py::ssize_t)rank);
        std::vector<int64_t> piv((size_t)rank);
        check(mpisdf_eri_cholesky_copy(h, factor.mutable_data(), piv.data(), residual.mutable_data()));
        for (int64_t k = 0; k < rank; ++k) pivots.mutable_data()[k] = (py::ssize_t)piv[k];
        return py::make_tuple(factor, pivots, residual);
    }, py::arg("context"), py::arg("tolerance") = 1e-12, py::arg("max_rank") = 0,
       py::arg("algorithm") = "unblocked_greedy", py::arg("debug") = false, py::arg("seed") = 0,
       "Return (factors[rank,npair], zero-based pivots, residual diagonal) in packed AO-pair order.");
    m.def("coulomb_metric", [cholesky_algorithm](const Context& ctx, const FloatArray& transform,
                                                 const std::string& algorithm, double tolerance, int64_t max_rank,
                                                 int64_t batch_size, bool debug, uint64_t seed) {
        const int32_t alg = cholesky_algorithm(algorithm);
        const int64_t nao = ctx.nao();
        if (transform.ndim() != 2 || transform.shape(1) != nao * nao)
            throw py::value_error("expected transform[nvec,nao*nao]");
        const int64_t nvec = transform.shape(0);
        py::array_t<double> result({(py::ssize_t)nvec, (py::ssize_t)nvec});
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_coulomb_metric(ctx.get(), transform.data(), nvec, alg, tolerance, max_rank, batch_size,
                                           debug ? 1 : 0, seed, result.mutable_data());
        }
        check(status);
        return result;
    }, py::arg("context"), py::arg("transform"), py::arg("algorithm") = "unblocked_greedy",
       py::arg("tolerance") = 1e-12, py::arg("max_rank") = 0, py::arg("batch_size") = 64,
       py::arg("debug") = false, py::arg("seed") = 0,
       "AO-product Coulomb metric; transform uses ordered AO-pair columns. No anchor weights applied.");

    m.def("isdf_jk", [](const FloatArray& x, const FloatArray& z, const FloatArray& dm) {
        if (x.ndim() != 2) throw py::value_error("expected anchor AO values X[k,nao]");
        const py::ssize_t k = x.shape(0), nao = x.shape(1);
        if (dm.ndim() != 2 || dm.shape(0) != nao || dm.shape(1) != nao)
            throw py::value_error("dm must be a finite AO density matrix of shape (nao,nao)");
        if (z.ndim() != 2 || z.shape(0) != k || z.shape(1) != k)
            throw py::value_error("Coulomb metric has the wrong anchor dimensions");
        py::array_t<double> j({nao, nao}), kmat({nao, nao});
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_isdf_jk(x.data(), k, nao, z.data(), dm.data(), j.mutable_data(), kmat.mutable_data());
        }
        check(status);
        return py::make_tuple(j, kmat);
    }, py::arg("x"), py::arg("metric"), py::arg("dm"),
       "ISDF J and K of a symmetric AO density from anchor AO values X and the anchor Coulomb metric.");

    // RBRP (cpp/src/anchor_select_pybind.cpp:68-140): the reference's result dict.
    auto rbrp_call = [](const Context* ctx, const FloatArray& data, const FloatArray& weights, int64_t nao,
                        int64_t batch_size, int64_t block_size, double tolerance, int64_t max_rank, uint64_t seed,
                        py::object tau_block, double rank_rtol, int64_t refresh_every, double svd_rtol,
                        bool compute_id) {
        const py::ssize_t n = data.shape(0);
        const double tb = tau_block.is_none() ? std::numeric_limits<double>::quiet_NaN() : tau_block.cast<double>();
        const int64_t p = nao * (nao + 1) / 2;
        const int64_t cap = std::max<int64_t>(1, std::min<int64_t>({p, (int64_t)n, max_rank ? max_rank : (int64_t)n}));
        std::vector<int64_t> idx(cap);
        std::vector<double> theta(compute_id ? (size_t)cap * n : 0), transform(compute_id ? (size_t)cap * nao * nao : 0),
            history((size_t)cap * 4), scalars(3);
        int64_t count = 0, nblocks = 0;
        int32_t status_code = 0;
        int status;
        {
            py::gil_scoped_release release;
            status = ctx ? mpisdf_rbrp_points(ctx->get(), data.data(), weights.data(), n, batch_size, block_size,
                                               tolerance, max_rank, seed, tb, rank_rtol, refresh_every, svd_rtol,
                                               compute_id ? 1 : 0, idx.data(), cap, &count,
                                               compute_id ? theta.data() : nullptr,
                                               compute_id ? transform.data() : nullptr, history.data(), &nblocks,
                                               scalars.data(), &status_code)
                         : mpisdf_rbrp_ao(data.data(), weights.data(), n, nao, batch_size, block_size, tolerance,
                                          max_rank, seed, tb, rank_rtol, refresh_every, svd_rtol, compute_id ? 1 : 0,
                                          idx.data(), cap, &count, compute_id ? theta.data() : nullptr,
                                          compute_id ? transform.data() : nullptr, history.data(), &nblocks,
                                          scalars.data(), &status_code);
        }
        check(status);
        static const char* names[] = {"rank_cap", "tolerance", "no_candidates", "numerical_rank", "roundoff_limited"};
        py::dict out;
        idx.resize(count);
        out["indices"] = py::cast(idx);
        out["relative_residual"] = scalars[0];
        out["status"] = std::string(names[status_code]);
        out["condition"] = scalars[1];
        out["retained_rank"] = (int64_t)scalars[2];
        py::array_t<double> hist({(py::ssize_t)nblocks, (py::ssize_t)4});
        std::copy(history.begin(), history.begin() + 4 * nblocks, hist.mutable_data());
        out["history"] = hist;
        if (compute_id) {
            py::array_t<double> th({(py::ssize_t)count, n}), tr({(py::ssize_t)count, (py::ssize_t)(nao * nao)});
            std::copy(theta.begin(), theta.begin() + (size_t)count * n, th.mutable_data());
            std::copy(transform.begin(), transform.begin() + (size_t)count * nao * nao, tr.mutable_data());
            out["interp_vecs"] = th;
            out["transform"] = tr;
        }
        return out;
    };
    m.def("rbrp_ao", [rbrp_call](const FloatArray& ao, const FloatArray& weights, int64_t batch_size,
                                 int64_t block_size, double tolerance, int64_t max_rank, uint64_t seed,
                                 py::object tau_block, double rank_rtol, int64_t refresh_every, double svd_rtol,
                                 bool compute_id) {
        if (ao.ndim() != 2 || weights.ndim() != 1 || weights.shape(0) != ao.shape(0))
            throw py::value_error("expected ao[grid,nao] and weights[grid]");
        return rbrp_call(nullptr, ao, weights, ao.shape(1), batch_size, block_size, tolerance, max_rank, seed,
                         tau_block, rank_rtol, refresh_every, svd_rtol, compute_id);
    }, py::arg("ao"), py::arg("weights"), py::arg("batch_size") = 128, py::arg("block_size") = 16,
       py::arg("tolerance") = 1e-10, py::arg("max_rank") = 0, py::arg("seed") = 7, py::arg("tau_block") = py::none(),
       py::arg("rank_rtol") = 1e-12, py::arg("refresh_every") = 8, py::arg("svd_rtol") = 1e-12,
       py::arg("compute_id") = false,
       "AO-only RBRP on nonzero grid rows with the optional Algorithm 5.1 ID (the reference's rbrp_ao_cpu "
       "surface; pair-row products on the GPU).");
    m.def("rbrp_ao_device", [rbrp_call](const Selector& selector, const FloatArray& points, const FloatArray& weights,
                                        int64_t batch_size, int64_t block_size, double tolerance, int64_t max_rank,
                                        uint64_t seed, py::object tau_block, double rank_rtol, int64_t refresh_every,
                                        double svd_rtol, bool compute_id) {
        if (points.ndim() != 2 || points.shape(1) != 3) throw py::value_error("expected points[ngrid,3]");
        if (weights.ndim() != 1 || weights.shape(0) != points.shape(0)) throw py::value_error("expected weights[ngrid]");
        return rbrp_call(&selector.context(), points, weights, selector.context().nao(), batch_size, block_size,
                         tolerance, max_rank, seed, tau_block, rank_rtol, refresh_every, svd_rtol, compute_id);
    }, py::arg("selector"), py::arg("points"), py::arg("weights"), py::arg("batch_size") = 128,
       py::arg("block_size") = 16, py::arg("tolerance") = 1e-10, py::arg("max_rank") = 0, py::arg("seed") = 7,
       py::arg("tau_block") = py::none(), py::arg("rank_rtol") = 1e-12, py::arg("refresh_every") = 8,
       py::arg("svd_rtol") = 1e-12, py::arg("compute_id") = false,
       "AO-only RBRP on grid points (AOs from the GPU kernel) with the optional ID.");

    m.def("ao_adjoint", [](const Context& ctx, const FloatArray& coords, const FloatArray& abar) {
        if (coords.ndim() != 2 || coords.shape(1) != 3) throw py::value_error("expected coords[n,3]");
        int64_t nao = 0;
        check(mpisdf_context_nao(ctx.get(), &nao));
        const py::ssize_t n = coords.shape(0);
        if (abar.ndim() != 2 || abar.shape(0) != n || abar.shape(1) != nao)
            throw py::value_error("expected abar[n,nao]");
        py::array_t<double> gout({(py::ssize_t)3, (py::ssize_t)nao});
        int status;
        {
            py::gil_scoped_release release;
            status = mpisdf_ao_adjoint_host(ctx.get(), coords.data(), n, abar.data(), gout.mutable_data());
        }
        check(status);
        return gout;
    }, py::arg("context"), py::arg("coords"), py::arg("abar"),
       "sum_p abar[p, mu] d phi_mu(r_p) / d r (shape [3, nao]), AO gradients recomputed on the GPU.");

    py::class_<Selector>(m, "AnchorSelection")
        .def(py::init<const Context&>(), py::arg("context"), py::keep_alive<1, 2>())
        .def("select", [](const Selector& sel, const FloatArray& points, const FloatArray& weights, py::object axes,
                          double tolerance, int64_t max_rank, LocalAlgorithm algorithm, py::object mo_coeff,
                          int64_t nocc, const std::string& space, int64_t batch_size, bool debug, int64_t sketch_size,
                          int64_t oversampling, uint64_t seed, bool audit, py::object cvt_weights, int64_t cvt_max_iter,
                          double cvt_tol, const std::vector<double>& axis_weights, py::object point_importance,
                          double ao_screening_threshold, const std::string& precision) {
            return sel.select(points, weights, axes, tolerance, max_rank, algorithm, mo_coeff, nocc, space, batch_size,
                              debug, sketch_size, oversampling, seed, audit, cvt_weights, cvt_max_iter, cvt_tol,
                              axis_weights, point_importance, ao_screening_threshold, precision);
        }, py::arg("points"), py::arg("weights"), py::arg("axes"),
             py::arg("tolerance") = 1e-10, py::arg("max_rank") = 0,
             py::arg("algorithm") = LocalAlgorithm::QRColumnPivoting, py::arg("mo_coeff") = py::none(),
             py::arg("nocc") = 0, py::arg("space") = "auto", py::arg("batch_size") = 128,
             py::arg("debug") = false, py::arg("sketch_size") = 128, py::arg("oversampling") = 16,
             py::arg("seed") = 7, py::arg("audit") = false, py::arg("cvt_weights") = py::none(),
             py::arg("cvt_max_iter") = 100, py::arg("cvt_tol") = 1e-5,
             py::arg("axis_weights") = std::vector<double>{}, py::arg("point_importance") = py::none(),
             py::arg("ao_screening_threshold") = 1e-24, py::arg("precision") = "fp64",
             "Select original grid indices in pivot order (GPU).  precision selects the sketch "
             "arithmetic of RandomizedQuadratureKhatriRao: fp64 (reference semantics), fp32, bf16 or fp8.")
        .def("select_with_diagnostics", &Selector::select_with_diagnostics, py::arg("points"), py::arg("weights"),
             py::arg("axes"), py::arg("tolerance") = 1e-10, py::arg("max_rank") = 0,
             py::arg("algorithm") = LocalAlgorithm::QRColumnPivoting, py::arg("mo_coeff") = py::none(),
             py::arg("nocc") = 0, py::arg("space") = "auto", py::arg("batch_size") = 128, py::arg("debug") = false,
             py::arg("sketch_size") = 128, py::arg("oversampling") = 16, py::arg("seed") = 7,
             py::arg("axis_weights") = std::vector<double>{}, py::arg("precision") = "fp64",
             py::arg("audit") = false, py::arg("point_importance") = py::none(),
             py::arg("ao_screening_threshold") = 1e-24,
             "select() plus a dict of stage timings and QRCP block statistics (not in the reference); with "
             "audit=True, 'audit' (and 'audit_density' for the density sketches) holds the projection residuals "
             "of the selected and of the dense QRCP pivots.")
        .def("evaluate_axis", &Selector::evaluate_axis, py::arg("points"), py::arg("axes"),
             py::arg("mo_coeff") = py::none(), py::arg("nocc") = 0, py::arg("space") = "auto",
             "Evaluate concatenated axis functions on the GPU.");

    struct Distributed {
        explicit Distributed(const Context&) {}
    };
    py::class_<Distributed>(m, "DistributedAnchorSelection")
        .def(py::init<const Context&>(), py::arg("context"), py::keep_alive<1, 2>())
        .def("select", [](const Distributed&, py::object, py::object, py::object, DistributedAlgorithm) {
            throw std::runtime_error("distributed SLATE SVD anchor selection is not implemented");
        }, py::arg("points"), py::arg("weights"), py::arg("axes"),
           py::arg("algorithm") = DistributedAlgorithm::SVD);
}
