// AO evaluation on grid tiles for gfx950.
//
// Computes what the reference kernel ``isdf_eval_ao`` computes
// (cpp/src/ao_eval.cpp:205-250): contracted real-spherical (or cartesian) AO
// values R(r) * sum_c T[s,c] x^lx y^ly z^lz and their grid-coordinate
// gradients R * dAng + 2 x R' * Ang, with R' = d R / d r^2.
//
// Layout and schedule (MI355X-first, not the reference's one work item per
// (point, shell) with stride-nao stores):
//   * one 256-thread workgroup owns a tile of 64 grid points;
//   * the tile's bounding sphere culls whole shells against per-shell radii
//     beyond which every value/derivative is below the precision's threshold;
//   * shells are visited in AO chunks; each chunk's [points x AOs] block is
//     staged in LDS and written back as contiguous rows of the grid-major
//     ``out[comp][point][ao]`` layout (coalesced full-row stores);
//   * angular parts are unrolled per angular momentum (template L), the
//     transform tables live in LDS.
#include "context.hpp"
#include "kernels.hpp"
#include "util.hpp"

#include <algorithm>

namespace mpisdf {
namespace {

constexpr int kTile = 64;       // grid points per workgroup
constexpr int kThreads = 256;   // 4 waves
constexpr int kTileBudget = 64 * 1024;
constexpr int kMaxChunkShells = 1024;

template <class C> __device__ __forceinline__ C dexp(C x);
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }

template <class T> __device__ __forceinline__ T convert_out(float v);
template <class T> __device__ __forceinline__ T convert_out(double v);
template <> __device__ __forceinline__ double convert_out<double>(double v) { return v; }
template <> __device__ __forceinline__ float convert_out<float>(float v) { return v; }
template <> __device__ __forceinline__ float convert_out<float>(double v) { return (float)v; }
template <> __device__ __forceinline__ __bf16 convert_out<__bf16>(float v) { return static_cast<__bf16>(v); }
template <> __device__ __forceinline__ __bf16 convert_out<__bf16>(double v) { return static_cast<__bf16>((float)v); }

// Evaluate one shell at one point into the LDS tile.  NC = output
// components: 1 values, 3 gradient, 4 values + gradient.
template <int L, class C, int NC>
__device__ __forceinline__ void eval_shell(const DevShell& sh, const double* __restrict__ env,
                                           const C* __restrict__ tab, C x, C y, C z, C r2,
                                           C* __restrict__ tile, int p, int col0, int width)
{
    constexpr int NCART = (L + 1) * (L + 2) / 2;
    C xp[L + 1], yp[L + 1], zp[L + 1];
    xp[0] = yp[0] = zp[0] = C(1);
#pragma unroll
    for (int i = 1; i <= L; ++i) { xp[i] = xp[i - 1] * x; yp[i] = yp[i - 1] * y; zp[i] = zp[i - 1] * z; }
    C mono[NCART], mx[NCART], my[NCART], mz[NCART];
    {
        int idx = 0;
#pragma unroll
        for (int lx = L; lx >= 0; --lx) {
#pragma unroll
            for (int ly = L - lx; ly >= 0; --ly) {
                const int lz = L - lx - ly;
                mono[idx] = xp[lx] * yp[ly] * zp[lz];
                if (NC > 1) {
                    mx[idx] = lx ? C(lx) * xp[lx - 1] * yp[ly] * zp[lz] : C(0);
                    my[idx] = ly ? C(ly) * xp[lx] * yp[ly - 1] * zp[lz] : C(0);
                    mz[idx] = lz ? C(lz) * xp[lx] * yp[ly] * zp[lz - 1] : C(0);
                }
                ++idx;
            }
        }
    }
    const int ncomp = sh.ncomp;
    for (int ctr = 0; ctr < sh.nctr; ++ctr) {
        C radial = 0, radial_prime = 0;
        for (int k = 0; k < sh.nprim; ++k) {
            const C alpha = (C)env[sh.exp_off + k];
            const C term = (C)env[sh.coeff_off + ctr * sh.nprim + k] * dexp<C>(-alpha * r2);
            radial += term;
            radial_prime -= alpha * term;
        }
        for (int s = 0; s < ncomp; ++s) {
            const C* t = tab + s * NCART;
            C value = 0, dx = 0, dy = 0, dz = 0;
#pragma unroll
            for (int c = 0; c < NCART; ++c) {
                value += t[c] * mono[c];
                if (NC > 1) { dx += t[c] * mx[c]; dy += t[c] * my[c]; dz += t[c] * mz[c]; }
            }
            const int col = col0 + ctr * ncomp + s;
            if (NC == 1) {
                tile[p * width + col] = radial * value;
            } else {
                const C gx = radial * dx + C(2) * x * radial_prime * value;
                const C gy = radial * dy + C(2) * y * radial_prime * value;
                const C gz = radial * dz + C(2) * z * radial_prime * value;
                const int plane = kTile * width;
                if (NC == 3) {
                    tile[p * width + col] = gx;
                    tile[plane + p * width + col] = gy;
                    tile[2 * plane + p * width + col] = gz;
                } else {
                    tile[p * width + col] = radial * value;
                    tile[plane + p * width + col] = gx;
                    tile[2 * plane + p * width + col] = gy;
                    tile[3 * plane + p * width + col] = gz;
                }
            }
        }
    }
}

// ADJOINT = true turns the kernel into the adjoint of gradient evaluation:
// nothing is stored; instead each tile writes its partial
// gout[tile * NC * nao + c * nao + mu] = sum_{p in tile} abar[p * ldab + mu]
// d_c phi_mu(r_p) (NC = 3), i.e. the AO gradients are recomputed tile by tile
// and contracted with the incoming adjoint without ever reaching HBM; the
// launcher adds the tiles' partials in tile order.
template <class C, class OutT, int NC, bool ADJOINT = false>
__global__ __launch_bounds__(kThreads) void ao_eval_kernel(
    const DevShell* __restrict__ shells, const double* __restrict__ env,
    const double* __restrict__ tabs, int ntab, const AoChunk* __restrict__ chunks, int nchunks,
    const double* __restrict__ xyz, int64_t n, OutT* __restrict__ out, int64_t ld, int64_t cstride,
    int width, const double* __restrict__ abar = nullptr, int64_t ldab = 0, double* __restrict__ gout = nullptr,
    int64_t nao = 0)
{
    extern __shared__ __align__(16) unsigned char smem[];
    C* tile = reinterpret_cast<C*>(smem);
    C* ltab = tile + NC * kTile * width;
    __shared__ double sx[kTile], sy[kTile], sz[kTile];
    __shared__ double centre[4];
    __shared__ int active[kMaxChunkShells];
    __shared__ int nactive;

    const int tid = threadIdx.x;
    const int64_t g0 = (int64_t)blockIdx.x * kTile;
    const int np = (int)min<int64_t>(kTile, n - g0);
    for (int i = tid; i < ntab; i += kThreads) ltab[i] = (C)tabs[i];
    if (tid < kTile) {
        const int64_t g = g0 + (tid < np ? tid : 0);
        sx[tid] = xyz[3 * g];
        sy[tid] = xyz[3 * g + 1];
        sz[tid] = xyz[3 * g + 2];
        // Bounding sphere of the tile: centroid, then max distance (wave 0).
        double cx = sx[tid], cy = sy[tid], cz = sz[tid];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            cx += __shfl_xor(cx, off);
            cy += __shfl_xor(cy, off);
            cz += __shfl_xor(cz, off);
        }
        cx *= 1.0 / kTile; cy *= 1.0 / kTile; cz *= 1.0 / kTile;
        const double dx = sx[tid] - cx, dy = sy[tid] - cy, dz = sz[tid] - cz;
        double rad = sqrt(dx * dx + dy * dy + dz * dz);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) rad = fmax(rad, __shfl_xor(rad, off));
        if (tid == 0) { centre[0] = cx; centre[1] = cy; centre[2] = cz; centre[3] = rad; }
    }
    __syncthreads();
    const double cx = centre[0], cy = centre[1], cz = centre[2], crad = centre[3];

    for (int c = 0; c < nchunks; ++c) {
        const AoChunk ck = chunks[c];
        const int cw = ck.a1 - ck.a0;
        for (int i = tid; i < NC * kTile * width; i += kThreads) tile[i] = C(0);
        if (tid == 0) nactive = 0;
        __syncthreads();
        for (int s = ck.s0 + tid; s < ck.s1; s += kThreads) {
            const DevShell& sh = shells[s];
            const double dx = sh.x - cx, dy = sh.y - cy, dz = sh.z - cz;
            const double rc = sqrt((double)(sizeof(C) == 8 ? sh.rcut2_f64 : sh.rcut2_f32));
            if (sqrt(dx * dx + dy * dy + dz * dz) - crad < rc) active[atomicAdd(&nactive, 1)] = s;
        }
        __syncthreads();
        const int items = nactive * kTile;
        for (int i = tid; i < items; i += kThreads) {
            const int p = i % kTile;
            if (p >= np) continue;
            const DevShell sh = shells[active[i / kTile]];
            const C x = (C)(sx[p] - sh.x), y = (C)(sy[p] - sh.y), z = (C)(sz[p] - sh.z);
            const C r2 = x * x + y * y + z * z;
            if ((double)r2 > (double)(sizeof(C) == 8 ? sh.rcut2_f64 : sh.rcut2_f32)) continue;
            const C* tab = ltab + sh.tab;
            const int col0 = sh.ao_off - ck.a0;
            switch (sh.l) {
            case 0: eval_shell<0, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            case 1: eval_shell<1, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            case 2: eval_shell<2, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            case 3: eval_shell<3, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            case 4: eval_shell<4, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            case 5: eval_shell<5, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            default: eval_shell<6, C, NC>(sh, env, tab, x, y, z, r2, tile, p, col0, width); break;
            }
        }
        __syncthreads();
        if (ADJOINT) {
            // Contract the staged gradient block with the adjoint rows and
            // reduce over the tile's points (adjoint rows read coalesced).
            for (int i = tid; i < NC * cw; i += kThreads) {
                const int comp = i / cw, j = i - comp * cw;
                const C* plane = tile + comp * kTile * width;
                const double* arow = abar + g0 * ldab + ck.a0 + j;
                double s = 0;
                for (int p = 0; p < np; ++p) s += (double)plane[p * width + j] * arow[(int64_t)p * ldab];
                gout[(int64_t)blockIdx.x * NC * nao + comp * nao + ck.a0 + j] = s;
            }
            __syncthreads();
            continue;
        }
        // Coalesced store of the staged block: each point's AO range is one
        // contiguous run in the grid-major output.
        for (int comp = 0; comp < NC; ++comp) {
            const C* plane = tile + comp * kTile * width;
            OutT* dst = out + comp * cstride + g0 * ld + ck.a0;
            for (int i = tid; i < np * cw; i += kThreads) {
                const int p = i / cw, j = i - p * cw;
                dst[(int64_t)p * ld + j] = convert_out<OutT>(plane[p * width + j]);
            }
        }
        __syncthreads();
    }
}

template <class C, class OutT, int NC>
void launch_typed(Context& ctx, const double* xyz, int64_t n, void* out, int64_t ld, int64_t cstride,
                  hipStream_t stream)
{
    int width = kTileBudget / (NC * kTile * (int)sizeof(C));
    width = std::max(width, ctx.max_shell_width);
    width = std::min<int64_t>(width, std::max<int64_t>(ctx.nao, 1));
    const auto& table = ctx.chunk_table(width);
    const int ntab = ctx.tab_offset[ctx.lmax + 1];
    const size_t lds = (size_t)NC * kTile * width * sizeof(C) + (size_t)ntab * sizeof(C);
    for (const auto& ck : table.first)
        if (ck.s1 - ck.s0 > kMaxChunkShells) throw std::logic_error("AO chunk holds too many shells");
    const int64_t blocks = (n + kTile - 1) / kTile;
    if (blocks > 0x7fffffff) throw std::overflow_error("AO grid is too large for one launch");
    auto kernel = ao_eval_kernel<C, OutT, NC>;
    if (lds > 64 * 1024)
        MPISDF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kThreads), lds, stream,
                       ctx.d_shells.ptr, ctx.d_env.ptr, ctx.d_tabs.ptr, ntab, table.second->ptr,
                       (int)table.first.size(), xyz, n, static_cast<OutT*>(out), ld, cstride, width,
                       static_cast<const double*>(nullptr), (int64_t)0, static_cast<double*>(nullptr), (int64_t)0);
    MPISDF_HIP(hipGetLastError());
}

template <class C, class OutT>
void launch_mode(Context& ctx, const double* xyz, int64_t n, AoMode mode, void* out, int64_t ld,
                 int64_t cstride, hipStream_t stream)
{
    switch (mode) {
    case kValues: launch_typed<C, OutT, 1>(ctx, xyz, n, out, ld, cstride, stream); break;
    case kGradient: launch_typed<C, OutT, 3>(ctx, xyz, n, out, ld, cstride, stream); break;
    case kValuesGradient: launch_typed<C, OutT, 4>(ctx, xyz, n, out, ld, cstride, stream); break;
    default: throw std::invalid_argument("AO mode must be 0 (values), 1 (gradient) or 2 (both)");
    }
}

}  // namespace

void launch_ao_adjoint(Context& ctx, const double* xyz, int64_t n, const double* abar, int64_t ldab, double* gout,
                       DeviceBuffer<double>& part, hipStream_t stream)
{
    if (n <= 0 || ctx.nao == 0) return;
    if (ldab < ctx.nao) throw std::invalid_argument("AO adjoint leading dimension is smaller than nao");
    constexpr int NC = 3;
    int width = kTileBudget / (NC * kTile * (int)sizeof(double));
    width = std::max(width, ctx.max_shell_width);
    width = std::min<int64_t>(width, std::max<int64_t>(ctx.nao, 1));
    const auto& table = ctx.chunk_table(width);
    const int ntab = ctx.tab_offset[ctx.lmax + 1];
    const size_t lds = (size_t)NC * kTile * width * sizeof(double) + (size_t)ntab * sizeof(double);
    for (const auto& ck : table.first)
        if (ck.s1 - ck.s0 > kMaxChunkShells) throw std::logic_error("AO chunk holds too many shells");
    auto kernel = ao_eval_kernel<double, double, NC, true>;
    if (lds > 64 * 1024)
        MPISDF_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // Launches of at most 4096 tiles, each tile's partial written to ``part``
    // and the tiles added in order, so the sum is the same on every run.
    constexpr int64_t kMaxTiles = 4096;
    const int64_t per = NC * ctx.nao;
    const int64_t tiles_total = (n + kTile - 1) / kTile;
    const int64_t need = std::min(tiles_total, kMaxTiles) * per;
    if (part.count < (size_t)need) part.resize((size_t)need);
    for (int64_t start = 0; start < n; start += kMaxTiles * kTile) {
        const int64_t cnt = std::min(n - start, kMaxTiles * kTile);
        const int64_t blocks = (cnt + kTile - 1) / kTile;
        hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(kThreads), lds, stream, ctx.d_shells.ptr, ctx.d_env.ptr,
                           ctx.d_tabs.ptr, ntab, table.second->ptr, (int)table.first.size(), xyz + 3 * start, cnt,
                           static_cast<double*>(nullptr), (int64_t)0, (int64_t)0, width, abar + start * ldab, ldab,
                           part.ptr, (int64_t)ctx.nao);
        MPISDF_HIP(hipGetLastError());
        launch_sum_partials(part.ptr, blocks, per, gout, true, stream);
    }
}

void launch_ao_eval(Context& ctx, const double* xyz, int64_t n, AoMode mode, DType dtype,
                    void* out, int64_t ld, int64_t cstride, hipStream_t stream)
{
    if (n <= 0 || ctx.nao == 0) return;
    if (ld < ctx.nao) throw std::invalid_argument("AO output leading dimension is smaller than nao");
    switch (dtype) {
    case kF64: launch_mode<double, double>(ctx, xyz, n, mode, out, ld, cstride, stream); break;
    case kF32: launch_mode<float, float>(ctx, xyz, n, mode, out, ld, cstride, stream); break;
    case kBF16: launch_mode<float, __bf16>(ctx, xyz, n, mode, out, ld, cstride, stream); break;
    default: throw std::invalid_argument("AO output dtype must be f64, f32 or bf16");
    }
}

// ---------------------------------------------------------------------------
// Axis assembly: features[g][col + j] = scale_a * {AO, dX, dY, dZ, occupied}
// (cpp/include/anchor_select_impl.hpp:62-77).  ``ao`` holds [4][n][nao]
// (values, then gradient components); ``occ`` holds [n][nocc].
// ---------------------------------------------------------------------------
__global__ void assemble_axis_kernel(const double* __restrict__ ao, const double* __restrict__ occ,
                                     int64_t n, int64_t nao, int64_t nocc, AxisSegment seg,
                                     double* __restrict__ features, int64_t width)
{
    const int64_t total = n * seg.width;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = idx / seg.width, j = idx - g * seg.width;
        double value;
        if (seg.kind == 4) value = occ[g * nocc + j];
        else value = ao[(int64_t)seg.kind * n * nao + g * nao + j];
        features[g * width + seg.offset + j] = seg.scale * value;
    }
}

void launch_assemble_axis(const double* ao, const double* occ, int64_t n, int64_t nao, int64_t nocc,
                          const AxisSegment& seg, double* features, int64_t width, hipStream_t stream)
{
    const int64_t total = n * seg.width;
    if (total <= 0) return;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(assemble_axis_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, ao, occ, n,
                       nao, nocc, seg, features, width);
    MPISDF_HIP(hipGetLastError());
}

}  // namespace mpisdf
