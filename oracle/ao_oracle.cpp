// CPU ORACLE -- test infrastructure, not product code.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker or as the timed CPU baseline.  The
// product (mp_isdf_amd) never links, imports or calls it.
//
// FP64 restatement of the reference's AO evaluator and axis assembly:
//   * shell table + cart->sph transform   cpp/src/ao_eval.cpp:78-158
//   * contracted AO values / gradients    cpp/src/ao_eval.cpp:164-252
//   * libcint cartesian order             cpp/src/ao_eval.cpp:30-44
//   * axis spec / feature concatenation   cpp/src/anchor_select.cpp:18-52,
//                                         cpp/include/anchor_select_impl.hpp:11-81
//   * multi-axis Frobenius normalisation  cpp/src/anchor_select.cpp:103-147
//
// The libcint CINTc2s_bra_sph table (libcint v6.2.1, a dependency absent from
// the reference tree) is restated from its published definition: real regular
// solid harmonics r^l Y_lm normalised on the unit sphere, m = -l..l for l >= 2,
// p in x,y,z order, with the l=0/1 factors 1/sqrt(4pi), sqrt(3/(4pi)) that
// the reference applies explicitly (ao_eval.cpp:135-137).  Coefficients come
// from the closed form of Helgaker, Jorgensen & Olsen, Molecular
// Electronic-Structure Theory, eq. 6.4.47-6.4.50.
//
// Parity status: pinned by closed forms (s/p shells), sphere orthonormality of
// the harmonics, radial normalisation integrals and central finite
// differences of the values (tests/test_oracle_ao.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;
constexpr int PTR_COORD = 1;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6;
constexpr int ANG_MAX = 6;
const double PI = 3.14159265358979323846;

thread_local std::string g_error;

struct Shell {
    int l, nprim, nctr, ncart, ncomp;
    int exp_off, coeff_off, ao_off;
    double x, y, z;
    std::vector<double> t;  // ncomp x ncart
};

double factorial(int n) { double r = 1; for (int i = 2; i <= n; ++i) r *= i; return r; }
double binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    return factorial(n) / (factorial(k) * factorial(n - k));
}

// Index of monomial (lx,ly,lz) in libcint's cartesian order for shell l.
int cart_index(int l, int lx, int ly) {
    // rows lx = l..0; within a row ly runs from l-lx down to 0.
    int idx = 0;
    for (int a = l; a > lx; --a) idx += l - a + 1;
    return idx + (l - lx - ly);
}

// Real solid harmonic S_lm (Racah normalisation) times sqrt((2l+1)/(4pi)),
// expanded over cartesian monomials of degree l.
std::vector<double> solid_harmonic_row(int l, int m) {
    const int am = std::abs(m);
    std::vector<double> row((l + 1) * (l + 2) / 2, 0.0);
    const double nlm = 1.0 / (std::pow(2.0, am) * factorial(l)) *
        std::sqrt(2.0 * factorial(l + am) * factorial(l - am) / (m == 0 ? 2.0 : 1.0));
    const double unit_sphere = std::sqrt((2 * l + 1) / (4 * PI));
    for (int t = 0; t <= (l - am) / 2; ++t) {
        for (int u = 0; u <= t; ++u) {
            for (int k = (m < 0 ? 1 : 0); k <= am; k += 2) {  // k = 2v
                const int sign_exp = t + (m < 0 ? (k - 1) / 2 : k / 2);
                const double c = ((sign_exp & 1) ? -1.0 : 1.0) * std::pow(0.25, t) *
                    binom(l, t) * binom(l - t, am + t) * binom(t, u) * binom(am, k);
                const int ly = 2 * u + k;
                const int lx = 2 * t + am - ly;
                const int lz = l - 2 * t - am;
                if (lx < 0 || lz < 0) continue;
                row[cart_index(l, lx, ly)] += c * nlm * unit_sphere;
            }
        }
    }
    return row;
}

std::vector<double> transform_for(int l, bool cart) {
    const int ncart = (l + 1) * (l + 2) / 2;
    const double sp = l == 0 ? 1.0 / std::sqrt(4 * PI) : l == 1 ? std::sqrt(3.0 / (4 * PI)) : 1.0;
    if (cart) {
        std::vector<double> t(ncart * ncart, 0.0);
        for (int c = 0; c < ncart; ++c) t[c * ncart + c] = sp;
        return t;
    }
    const int nsph = 2 * l + 1;
    std::vector<double> t(nsph * ncart, 0.0);
    if (l == 1) {  // libcint keeps p functions in x, y, z order
        for (int c = 0; c < 3; ++c) t[c * 3 + c] = sp;
        return t;
    }
    for (int s = 0; s < nsph; ++s) {
        const auto row = solid_harmonic_row(l, s - l);
        std::copy(row.begin(), row.end(), t.begin() + s * ncart);
    }
    return t;
}

void cart_powers(int l, int index, int& lx, int& ly, int& lz) {
    for (lx = l; lx >= 0; --lx) {
        const int count = l - lx + 1;
        if (index < count) { ly = l - lx - index; lz = index; return; }
        index -= count;
    }
    lx = ly = lz = 0;
}

double ipow(double x, int n) { double v = 1; for (int i = 0; i < n; ++i) v *= x; return v; }

std::vector<Shell> build_shells(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                                const double* env, int64_t nenv, bool cart, int64_t& nao)
{
    std::vector<Shell> shells;
    nao = 0;
    auto check = [&](int64_t off, int64_t len) {
        if (off < 0 || len < 0 || off + len > nenv)
            throw std::invalid_argument("libcint table contains an invalid ENV offset");
    };
    for (int i = 0; i < nbas; ++i) {
        const int32_t* b = bas + i * BAS_SLOTS;
        const int atom = b[ATOM_OF], l = b[ANG_OF], np = b[NPRIM_OF], nc = b[NCTR_OF];
        if (atom < 0 || atom >= natm || l < 0 || l > ANG_MAX || np <= 0 || nc <= 0)
            throw std::invalid_argument("invalid libcint basis shell");
        check(b[PTR_EXP], np);
        check(b[PTR_COEFF], (int64_t)np * nc);
        const int coord = atm[atom * ATM_SLOTS + PTR_COORD];
        check(coord, 3);
        Shell s;
        s.l = l; s.nprim = np; s.nctr = nc;
        s.ncart = (l + 1) * (l + 2) / 2;
        s.ncomp = cart ? s.ncart : 2 * l + 1;
        s.exp_off = b[PTR_EXP]; s.coeff_off = b[PTR_COEFF];
        s.ao_off = (int)nao;
        s.x = env[coord]; s.y = env[coord + 1]; s.z = env[coord + 2];
        s.t = transform_for(l, cart);
        nao += (int64_t)s.ncomp * nc;
        shells.push_back(std::move(s));
    }
    return shells;
}

// deriv: 0 -> values [n][nao]; 1 -> gradient [3][n][nao]; 2 -> values+gradient [4][n][nao]
void ao_eval(const std::vector<Shell>& shells, const double* env, int64_t nao,
             const double* xyz, int64_t n, int deriv, double* out)
{
    const int64_t stride = n * nao;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < n; ++p) {
        for (const Shell& b : shells) {
            const double x = xyz[3 * p] - b.x, y = xyz[3 * p + 1] - b.y, z = xyz[3 * p + 2] - b.z;
            const double r2 = x * x + y * y + z * z;
            for (int ctr = 0; ctr < b.nctr; ++ctr) {
                double radial = 0, radial_prime = 0;
                for (int k = 0; k < b.nprim; ++k) {
                    const double alpha = env[b.exp_off + k];
                    const double term = env[b.coeff_off + ctr * b.nprim + k] * std::exp(-alpha * r2);
                    radial += term;
                    radial_prime -= alpha * term;
                }
                for (int s = 0; s < b.ncomp; ++s) {
                    double value = 0, dx = 0, dy = 0, dz = 0;
                    for (int c = 0; c < b.ncart; ++c) {
                        int lx, ly, lz;
                        cart_powers(b.l, c, lx, ly, lz);
                        const double px = ipow(x, lx), py = ipow(y, ly), pz = ipow(z, lz);
                        const double t = b.t[s * b.ncart + c];
                        value += t * px * py * pz;
                        if (lx) dx += t * lx * ipow(x, lx - 1) * py * pz;
                        if (ly) dy += t * ly * px * ipow(y, ly - 1) * pz;
                        if (lz) dz += t * lz * px * py * ipow(z, lz - 1);
                    }
                    const int64_t index = p * nao + b.ao_off + ctr * b.ncomp + s;
                    const double gx = radial * dx + 2.0 * x * radial_prime * value;
                    const double gy = radial * dy + 2.0 * y * radial_prime * value;
                    const double gz = radial * dz + 2.0 * z * radial_prime * value;
                    if (deriv == 0) {
                        out[index] = radial * value;
                    } else if (deriv == 1) {
                        out[index] = gx; out[stride + index] = gy; out[2 * stride + index] = gz;
                    } else {
                        out[index] = radial * value;
                        out[stride + index] = gx; out[2 * stride + index] = gy;
                        out[3 * stride + index] = gz;
                    }
                }
            }
        }
    }
}

template <class F>
int guarded(F&& f)
{
    try { f(); return 0; }
    catch (const std::invalid_argument& e) { g_error = e.what(); return 1; }
    catch (const std::overflow_error& e) { g_error = e.what(); return 2; }
    catch (const std::exception& e) { g_error = e.what(); return 3; }
}

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_error.c_str(); }

// Transform (c2s) table of one angular momentum: ncomp x ncart, row-major.
int oracle_c2s(int l, int cart, double* out)
{
    return guarded([&] {
        if (l < 0 || l > ANG_MAX) throw std::invalid_argument("l out of range");
        const auto t = transform_for(l, cart != 0);
        std::copy(t.begin(), t.end(), out);
    });
}

int oracle_nao(const int32_t* atm, int natm, const int32_t* bas, int nbas,
               const double* env, int64_t nenv, int cart, int64_t* nao)
{
    return guarded([&] { build_shells(atm, natm, bas, nbas, env, nenv, cart != 0, *nao); });
}

int oracle_ao_eval(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                   const double* env, int64_t nenv, int cart,
                   const double* xyz, int64_t n, int deriv, double* out)
{
    return guarded([&] {
        int64_t nao = 0;
        const auto shells = build_shells(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        if (deriv < 0 || deriv > 2) throw std::invalid_argument("deriv must be 0, 1 or 2");
        ao_eval(shells, env, nao, xyz, n, deriv, out);
    });
}

// Axis codes follow the reference AxisType order (cpp/include/enums.hpp:8-14):
// 0 AO, 1 dX, 2 dY, 3 dZ, 4 occupied.  out is [n][width] row-major with the
// scales applied per axis (anchor_select_impl.hpp:62-77).  mo_coeff is the
// row-major (nao, nao) matrix whose first nocc columns form the occupied axis.
int oracle_axis_eval(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                     const double* env, int64_t nenv, int cart,
                     const double* xyz, int64_t n, const int32_t* axes, int naxes,
                     const double* mo_coeff, int64_t nocc, const double* scales,
                     double* out, int64_t* width_out)
{
    return guarded([&] {
        int64_t nao = 0;
        const auto shells = build_shells(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        if (naxes <= 0) throw std::invalid_argument("axis list must not be empty");
        int64_t width = 0;
        bool need_occ = false;
        for (int a = 0; a < naxes; ++a) {
            if (axes[a] < 0 || axes[a] > 4) throw std::invalid_argument("unknown AxisType");
            if (axes[a] == 4) need_occ = true;
            width += axes[a] == 4 ? nocc : nao;
        }
        if (need_occ && (nocc <= 0 || nocc > nao || !mo_coeff))
            throw std::invalid_argument("OccupiedOrbital requires mo_coeff and 1 <= nocc <= nao");
        *width_out = width;
        if (!out) return;
        std::vector<double> both((size_t)(4 * n * nao));
        ao_eval(shells, env, nao, xyz, n, 2, both.data());
        int64_t col = 0;
        for (int a = 0; a < naxes; ++a) {
            const double scale = scales ? scales[a] : 1.0;
            const int64_t w = axes[a] == 4 ? nocc : nao;
            for (int64_t g = 0; g < n; ++g) {
                for (int64_t j = 0; j < w; ++j) {
                    double value = 0;
                    if (axes[a] == 0) value = both[g * nao + j];
                    else if (axes[a] == 4) {
                        for (int64_t k = 0; k < nao; ++k) value += both[g * nao + k] * mo_coeff[k * nao + j];
                    } else value = both[axes[a] * n * nao + g * nao + j];
                    out[g * width + col + j] = scale * value;
                }
            }
            col += w;
        }
    });
}

// Global per-axis Frobenius norms over nonzero-weight points, accumulated with
// hypot in 128-point batches (cpp/src/anchor_select.cpp:103-147).  Writes
// scale_a = axis_weight_a / norm_a.
int oracle_axis_scales(const int32_t* atm, int natm, const int32_t* bas, int nbas,
                       const double* env, int64_t nenv, int cart,
                       const double* xyz, const double* weights, int64_t n,
                       const int32_t* axes, int naxes, const double* mo_coeff, int64_t nocc,
                       const double* axis_weights, double* scales_out)
{
    return guarded([&] {
        int64_t width = 0;
        if (oracle_axis_eval(atm, natm, bas, nbas, env, nenv, cart, xyz, 0, axes, naxes,
                             mo_coeff, nocc, nullptr, nullptr, &width))
            throw std::invalid_argument(g_error);
        int64_t nao = 0;
        build_shells(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        std::vector<double> norms(naxes, 0.0);
        std::vector<double> batch, values;
        bool any = false;
        auto accumulate = [&]() {
            const int64_t m = (int64_t)batch.size() / 3;
            if (!m) return;
            values.assign((size_t)(m * width), 0.0);
            int64_t w2 = 0;
            if (oracle_axis_eval(atm, natm, bas, nbas, env, nenv, cart, batch.data(), m, axes, naxes,
                                 mo_coeff, nocc, nullptr, values.data(), &w2))
                throw std::invalid_argument(g_error);
            for (int64_t g = 0; g < m; ++g) {
                int64_t off = 0;
                for (int a = 0; a < naxes; ++a) {
                    const int64_t w = axes[a] == 4 ? nocc : nao;
                    for (int64_t j = 0; j < w; ++j) norms[a] = std::hypot(norms[a], values[g * width + off + j]);
                    off += w;
                }
            }
            batch.clear();
        };
        for (int64_t g = 0; g < n; ++g) {
            if (!std::isfinite(weights[g])) throw std::invalid_argument("grid weights must be finite");
            if (weights[g] == 0) continue;
            any = true;
            batch.insert(batch.end(), xyz + 3 * g, xyz + 3 * g + 3);
            if ((int64_t)batch.size() == 3 * 128) accumulate();
        }
        accumulate();
        for (int a = 0; a < naxes; ++a) {
            if (!any) { scales_out[a] = 1.0; continue; }
            if (!(norms[a] > 0) || !std::isfinite(norms[a]))
                throw std::invalid_argument("axis Frobenius norm must be finite and positive");
            scales_out[a] = (axis_weights ? axis_weights[a] : 1.0) / norms[a];
        }
    });
}

}  // extern "C"
