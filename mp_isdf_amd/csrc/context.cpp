// Basis context construction (host side).
//
// Follows the reference constructor cpp/src/ao_eval.cpp:78-158: validate the
// libcint ATM/BAS/ENV tables, record per-shell (l, nprim, nctr, offsets,
// centre, AO offset) and build the cart->sph transform.  The transform is
// built here from associated Legendre polynomials rather than calling
// libcint's CINTc2s_bra_sph: r^l P_l^m(z/r) e^{i m phi} = (x+iy)^m *
// sum_k a_k z^(l-m-2k) r^(2k), normalised on the unit sphere without the
// Condon-Shortley phase, m = -l..l for l >= 2 and x,y,z for p shells; the
// l = 0/1 normalisations equal the reference's explicit overrides
// (ao_eval.cpp:135-137).  Cartesian shells (no reference counterpart) use
// libcint's cartesian convention: monomials scaled by 1/sqrt(4pi) (s) and
// sqrt(3/(4pi)) (p), unscaled for l >= 2.
//
// Added for the GPU: per-shell culling radii (for FP64 and FP32 thresholds)
// and AO chunk tables that let one workgroup stage a contiguous AO range of a
// point tile in LDS before a coalesced store.
#include "context.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace mpisdf {
namespace {

constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;
constexpr int PTR_COORD = 1;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6;
constexpr double kPi = 3.14159265358979323846;
// Absolute culling thresholds on |value| and |d value / dr| of one AO.
constexpr double kCullF64 = 1e-18;
constexpr double kCullF32 = 1e-10;

double fact(int n) { double r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }

// Dense polynomial in x, y, z of total degree <= L, indexed [lx][ly][lz].
struct Poly {
    int L;
    std::vector<double> c;
    explicit Poly(int degree) : L(degree), c((degree + 1) * (degree + 1) * (degree + 1), 0.0) {}
    double& at(int a, int b, int d) { return c[(a * (L + 1) + b) * (L + 1) + d]; }
    double at(int a, int b, int d) const { return c[(a * (L + 1) + b) * (L + 1) + d]; }
};

Poly mul(const Poly& p, const Poly& q, int L)
{
    Poly r(L);
    for (int a = 0; a <= p.L; ++a)
        for (int b = 0; a + b <= p.L; ++b)
            for (int d = 0; a + b + d <= p.L; ++d) {
                const double pv = p.at(a, b, d);
                if (pv == 0) continue;
                for (int e = 0; e <= q.L; ++e)
                    for (int f = 0; e + f <= q.L; ++f)
                        for (int g = 0; e + f + g <= q.L; ++g) {
                            const double qv = q.at(e, f, g);
                            if (qv == 0 || a + e + b + f + d + g > L) continue;
                            r.at(a + e, b + f, d + g) += pv * qv;
                        }
            }
    return r;
}

// Row of the unit-sphere real solid harmonic (l, m) over libcint cartesian order.
std::vector<double> harmonic_row(int l, int m)
{
    const int am = std::abs(m);
    // m-th derivative of the Legendre polynomial P_l: coefficients of t^(l-am-2k).
    std::vector<double> a;
    for (int k = 0; 2 * k <= l - am; ++k) {
        const int power = l - 2 * k;  // power in P_l before differentiation
        const double pl = ((k & 1) ? -1.0 : 1.0) * fact(2 * l - 2 * k) /
                          (std::pow(2.0, l) * fact(k) * fact(l - k) * fact(l - 2 * k));
        a.push_back(pl * fact(power) / fact(power - am));
    }
    // Q = sum_k a_k z^(l-am-2k) (x^2+y^2+z^2)^k
    Poly q(l);
    Poly r2(2);
    r2.at(2, 0, 0) = r2.at(0, 2, 0) = r2.at(0, 0, 2) = 1.0;
    for (int k = 0; k < (int)a.size(); ++k) {
        Poly term(l);
        term.at(0, 0, l - am - 2 * k) = a[k];
        for (int i = 0; i < k; ++i) term = mul(term, r2, l);
        for (size_t i = 0; i < q.c.size(); ++i) q.c[i] += term.c[i];
    }
    // (x + i y)^am
    Poly re(l), im(l);
    re.at(0, 0, 0) = 1.0;
    for (int i = 0; i < am; ++i) {
        Poly nre(l), nim(l);
        for (int x = 0; x < l; ++x)
            for (int y = 0; x + y < l; ++y) {
                nre.at(x + 1, y, 0) += re.at(x, y, 0);
                nre.at(x, y + 1, 0) -= im.at(x, y, 0);
                nim.at(x, y + 1, 0) += re.at(x, y, 0);
                nim.at(x + 1, y, 0) += im.at(x, y, 0);
            }
        re = nre;
        im = nim;
    }
    const Poly s = mul(m >= 0 ? re : im, q, l);
    const double norm = std::sqrt((2 * l + 1) / (4 * kPi) * fact(l - am) / fact(l + am)) *
                        (m == 0 ? 1.0 : std::sqrt(2.0));
    std::vector<double> row;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) row.push_back(norm * s.at(lx, ly, l - lx - ly));
    return row;
}

double culling_radius(int l, const double* exps, const double* cmax, int nprim, double amax,
                      double threshold)
{
    double start = 0;
    double amin = std::numeric_limits<double>::max();
    for (int p = 0; p < nprim; ++p) amin = std::min(amin, exps[p]);
    start = std::sqrt((l + 1) / (2 * amin));
    auto bound = [&](double r) {
        double b = 0;
        for (int p = 0; p < nprim; ++p) {
            const double e = cmax[p] * std::exp(-exps[p] * r * r);
            b += e * amax * (std::pow(r, l) + (l ? l * std::pow(r, l - 1) : 0.0) +
                             2 * exps[p] * std::pow(r, l + 1));
        }
        return b;
    };
    for (double r = start; r < 200.0; r += 0.02)
        if (bound(r) < threshold) return r;
    return 200.0;
}

}  // namespace

std::vector<double> transform_table(int l, bool cart)
{
    const int ncart = (l + 1) * (l + 2) / 2;
    const double sp = l == 0 ? 1.0 / std::sqrt(4 * kPi) : l == 1 ? std::sqrt(3.0 / (4 * kPi)) : 1.0;
    if (cart || l <= 1) {
        std::vector<double> t(ncart * ncart, 0.0);
        for (int c = 0; c < ncart; ++c) t[c * ncart + c] = sp;
        return t;  // s: 1x1; p: x, y, z identity as libcint orders them
    }
    std::vector<double> t;
    for (int m = -l; m <= l; ++m) {
        const auto row = harmonic_row(l, m);
        t.insert(t.end(), row.begin(), row.end());
    }
    return t;
}

void* Context::reserve_sketch(size_t bytes)
{
    if (sketch_store.count < bytes) {
        sketch_store.release();  // before the new allocation: both may not fit
        sketch_store.resize(bytes);
    }
    return sketch_store.ptr;
}

void Context::release_sketch_store()
{
    std::lock_guard<std::mutex> lock(sketch_mutex);
    sketch_store.release();
}

const std::pair<std::vector<AoChunk>, std::unique_ptr<DeviceBuffer<AoChunk>>>&
Context::chunk_table(int width)
{
    std::lock_guard<std::mutex> lock(chunk_mutex);
    auto it = chunks.find(width);
    if (it != chunks.end()) return it->second;
    if (width < max_shell_width) throw std::logic_error("AO chunk narrower than a shell");
    std::vector<AoChunk> list;
    int s = 0;
    while (s < nshells) {
        AoChunk c{s, s, shells[s].ao_off, shells[s].ao_off};
        while (s < nshells) {
            const int w = shells[s].nctr * shells[s].ncomp;
            if (c.a1 - c.a0 + w > width) break;
            c.a1 += w;
            c.s1 = ++s;
        }
        list.push_back(c);
    }
    auto dev = std::make_unique<DeviceBuffer<AoChunk>>();
    dev->upload(list.data(), list.size(), nullptr);
    MPISDF_HIP(hipStreamSynchronize(nullptr));
    auto& slot = chunks[width];
    slot.first = std::move(list);
    slot.second = std::move(dev);
    return slot;
}

std::unique_ptr<Context> make_context(const int32_t* atm, int64_t natm, const int32_t* bas,
                                      int64_t nbas, const double* env, int64_t nenv, bool cart)
{
    if (natm < 0 || nbas < 0 || nenv < 0 || natm > std::numeric_limits<int>::max() ||
        nbas > std::numeric_limits<int>::max())
        throw std::invalid_argument("invalid libcint ATM/BAS layout");
    auto ctx = std::make_unique<Context>();
    ctx->atm.assign(atm, atm + natm * ATM_SLOTS);
    ctx->bas.assign(bas, bas + nbas * BAS_SLOTS);
    ctx->env.assign(env, env + nenv);
    ctx->cart = cart;
    auto check_range = [&](int64_t off, int64_t len) {
        if (off < 0 || len < 0 || off > nenv || len > nenv - off)
            throw std::invalid_argument("libcint table contains an invalid ENV offset");
    };
    int64_t nao = 0;
    for (int64_t i = 0; i < nbas; ++i) {
        const int32_t* b = ctx->bas.data() + i * BAS_SLOTS;
        const int atom = b[ATOM_OF], l = b[ANG_OF], np = b[NPRIM_OF], nc = b[NCTR_OF];
        if (atom < 0 || atom >= natm || l < 0 || l > kAngMax || np <= 0 || nc <= 0 ||
            np > 64 || nc > 64)
            throw std::invalid_argument("invalid libcint basis shell");
        check_range(b[PTR_EXP], np);
        check_range(b[PTR_COEFF], (int64_t)np * nc);
        const int coord = ctx->atm[atom * ATM_SLOTS + PTR_COORD];
        check_range(coord, 3);
        const int ncart = (l + 1) * (l + 2) / 2;
        const int ncomp = cart ? ncart : 2 * l + 1;
        if (nao + (int64_t)ncomp * nc > std::numeric_limits<int>::max())
            throw std::overflow_error("AO or transformation index exceeds int range");
        DevShell s{};
        s.x = env[coord];
        s.y = env[coord + 1];
        s.z = env[coord + 2];
        s.l = l;
        s.nprim = np;
        s.nctr = nc;
        s.ncomp = ncomp;
        s.exp_off = b[PTR_EXP];
        s.coeff_off = b[PTR_COEFF];
        s.ao_off = (int)nao;
        s.tab = l;
        ctx->shells.push_back(s);
        ctx->lmax = std::max(ctx->lmax, l);
        ctx->max_shell_width = std::max(ctx->max_shell_width, ncomp * nc);
        nao += (int64_t)ncomp * nc;
    }
    ctx->nao = nao;
    ctx->nshells = (int)ctx->shells.size();
    // Per-l transforms and culling radii.
    ctx->tab_offset.assign(kAngMax + 2, 0);
    std::vector<double> amax(kAngMax + 1, 0.0);
    for (int l = 0; l <= kAngMax; ++l) {
        ctx->tab_offset[l] = (int)ctx->tabs.size();
        const auto t = transform_table(l, cart);
        const int ncart = (l + 1) * (l + 2) / 2;
        for (size_t r = 0; r < t.size() / ncart; ++r) {
            double sum = 0;
            for (int c = 0; c < ncart; ++c) sum += std::abs(t[r * ncart + c]);
            amax[l] = std::max(amax[l], sum);
        }
        ctx->tabs.insert(ctx->tabs.end(), t.begin(), t.end());
    }
    ctx->tab_offset[kAngMax + 1] = (int)ctx->tabs.size();
    for (auto& s : ctx->shells) {
        std::vector<double> cmax(s.nprim, 0.0);
        for (int p = 0; p < s.nprim; ++p)
            for (int c = 0; c < s.nctr; ++c)
                cmax[p] = std::max(cmax[p], std::abs(env[s.coeff_off + c * s.nprim + p]));
        const double r64 = culling_radius(s.l, env + s.exp_off, cmax.data(), s.nprim, amax[s.l], kCullF64);
        const double r32 = culling_radius(s.l, env + s.exp_off, cmax.data(), s.nprim, amax[s.l], kCullF32);
        s.rcut2_f64 = (float)(r64 * r64);
        s.rcut2_f32 = (float)(r32 * r32);
        s.tab = ctx->tab_offset[s.l];
    }
    ensure_device();
    ctx->d_shells.upload(ctx->shells.data(), ctx->shells.size(), nullptr);
    ctx->d_env.upload(ctx->env.data(), ctx->env.size(), nullptr);
    ctx->d_tabs.upload(ctx->tabs.data(), ctx->tabs.size(), nullptr);
    MPISDF_HIP(hipStreamSynchronize(nullptr));
    return ctx;
}

}  // namespace mpisdf
