#include "axis.hpp"

#include <cmath>
#include <limits>

namespace mpisdf {

AxisSpec axis_spec(const Context& ctx, const int32_t* axes, int naxes, const double* mo_coeff,
                   int64_t nocc)
{
    if (naxes <= 0 || !axes) throw std::invalid_argument("axis list must not be empty");
    AxisSpec spec;
    for (int a = 0; a < naxes; ++a) {
        int64_t w = 0;
        switch (axes[a]) {
        case 0: spec.need_ao = true; w = ctx.nao; break;
        case 1: case 2: case 3: spec.need_grad = true; w = ctx.nao; break;
        case 4: spec.need_occ = true; w = nocc; break;
        default: throw std::invalid_argument("unknown AxisType");
        }
        if (w > std::numeric_limits<int64_t>::max() - spec.width)
            throw std::overflow_error("axis feature count is too large");
        spec.kinds.push_back(axes[a]);
        spec.widths.push_back(w);
        spec.width += w;
    }
    if (spec.need_occ) {
        if (ctx.nao == 0 || nocc <= 0 || nocc > ctx.nao || !mo_coeff)
            throw std::invalid_argument(
                "OccupiedOrbital requires row-major (nao,nao) RHF mo_coeff and 1 <= nocc <= nao");
        spec.need_ao = true;
    }
    if (spec.width == 0) throw std::invalid_argument("no orbital features available");
    return spec;
}

void evaluate_axis_device(Context& ctx, const AxisSpec& spec, const double* xyz, int64_t n,
                          const double* scales, const double* mo_host, int64_t nocc,
                          double* features, AxisWorkspace& ws, hipStream_t stream)
{
    if (n <= 0) return;
    const int64_t nao = ctx.nao;
    if (scales)
        for (size_t a = 0; a < spec.kinds.size(); ++a)
            if (!std::isfinite(scales[a]) || scales[a] <= 0)
                throw std::invalid_argument("axis scales must be finite and positive");
    if ((size_t)(4 * n * nao) > ws.ao.count) ws.ao.resize((size_t)(4 * n * nao));
    // One fused pass yields values and gradients when both are needed.
    const AoMode mode = spec.need_grad ? (spec.need_ao ? kValuesGradient : kGradient) : kValues;
    const int64_t cstride = n * nao;
    // Keep component index = axis kind: values at plane 0, d/dx..d/dz at 1..3.
    double* base = mode == kGradient ? ws.ao.ptr + cstride : ws.ao.ptr;
    launch_ao_eval(ctx, xyz, n, mode, kF64, base, nao, cstride, stream);
    if (spec.need_occ) {
        if (!ws.mo_ready) {
            ws.mo.upload(mo_host, (size_t)(nao * nao), stream);
            ws.mo_ready = true;
        }
        if ((size_t)(n * nocc) > ws.occ.count) ws.occ.resize((size_t)(n * nocc));
        gemm_f64(n, nocc, nao, 1.0, MatView{ws.ao.ptr, nao, 1}, MatView{ws.mo.ptr, nao, 1}, 0.0,
                 MutView{ws.occ.ptr, nocc, 1}, stream);
    }
    int64_t offset = 0;
    for (size_t a = 0; a < spec.kinds.size(); ++a) {
        AxisSegment seg{spec.kinds[a], offset, spec.widths[a], scales ? scales[a] : 1.0};
        launch_assemble_axis(ws.ao.ptr, ws.occ.ptr, n, nao, nocc, seg, features, spec.width, stream);
        offset += spec.widths[a];
    }
}

}  // namespace mpisdf
