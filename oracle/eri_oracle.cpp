// CPU ORACLE -- test infrastructure, not product code.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker or as the timed CPU baseline.  The
// product (mp_isdf_amd) never links, imports or calls it.
//
// FP64 restatement of the reference's integral-direct Cholesky Coulomb metric:
//   * shell-quartet ERIs (ij|kl) over the context's AOs: the reference calls
//     libcint's int2e (libcint v6.2.1, a dependency absent from the reference
//     tree).  Restated here from the published McMurchie-Davidson scheme
//     (Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure Theory,
//     eqs. 9.5.6-9.5.7 for the Hermite coefficients E, 9.9.18-9.9.20 for the
//     Hermite integrals R, 9.9.33 for the ERI) with the Boys function F_n
//     (its series, downward recursion and the erf form), contracted and
//     transformed with the oracle's own c2s tables (oracle_c2s).
//   * the greedy / uniform pivoted Cholesky     cpp/src/cholesky_eri_greedy.cpp:53-175
//   * pivot choice                              cpp/include/cholesky_pivot.hpp:16-36
//   * column correction (sequential k order)    cpp/include/cholesky_update.hpp:43-90
//   * the metric T V T^T through the factors    cpp/src/coulomb_metric.cpp:20-98
//
// Parity status: pinned by the closed-form s-Gaussian ERIs (Boys 1950,
// Szabo-Ostlund's H2/STO-3G values, through oracle.s_eri), by the
// translational-derivative relations that make p and d primitives out of s
// primitives, checked with high-order central differences, by the 8-fold
// permutational symmetry and positive semidefiniteness of the packed matrix,
// and by a NumPy restatement of the greedy loop on the dense packed matrix
// (the reference test's own reference_greedy, test_coulomb_metric.py:62-78)
// (tests/test_eri_oracle.py).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" int oracle_c2s(int l, int cart, double* out);

namespace {

constexpr int ATM_SLOTS = 6, BAS_SLOTS = 8;
constexpr int PTR_COORD = 1;
constexpr int ATOM_OF = 0, ANG_OF = 1, NPRIM_OF = 2, NCTR_OF = 3, PTR_EXP = 5, PTR_COEFF = 6;
constexpr int LMAX = 4;
const double PI = 3.14159265358979323846;

thread_local std::string g_error;

struct EShell {
    int l, nprim, nctr, ncart, ncomp, ao_off;
    double c[3];
    const double* exps;
    const double* coef;      // [nctr][nprim]
    std::vector<double> t;   // ncomp x ncart
    std::vector<int> pw;     // ncart x 3 cartesian powers (libcint order)
};

std::vector<EShell> shells_of(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env,
                              int64_t nenv, bool cart, int64_t& nao)
{
    std::vector<EShell> out;
    nao = 0;
    for (int i = 0; i < nbas; ++i) {
        const int32_t* b = bas + i * BAS_SLOTS;
        const int atom = b[ATOM_OF], l = b[ANG_OF], np = b[NPRIM_OF], nc = b[NCTR_OF];
        if (atom < 0 || atom >= natm || l < 0 || l > LMAX || np <= 0 || nc <= 0)
            throw std::invalid_argument("invalid libcint basis shell");
        const int coord = atm[atom * ATM_SLOTS + PTR_COORD];
        if (b[PTR_EXP] < 0 || b[PTR_EXP] + np > nenv || b[PTR_COEFF] < 0 || b[PTR_COEFF] + np * nc > nenv ||
            coord < 0 || coord + 3 > nenv)
            throw std::invalid_argument("libcint table contains an invalid ENV offset");
        EShell s;
        s.l = l;
        s.nprim = np;
        s.nctr = nc;
        s.ncart = (l + 1) * (l + 2) / 2;
        s.ncomp = cart ? s.ncart : 2 * l + 1;
        s.ao_off = (int)nao;
        s.c[0] = env[coord];
        s.c[1] = env[coord + 1];
        s.c[2] = env[coord + 2];
        s.exps = env + b[PTR_EXP];
        s.coef = env + b[PTR_COEFF];
        s.t.resize((size_t)s.ncomp * s.ncart);
        oracle_c2s(l, cart ? 1 : 0, s.t.data());
        for (int lx = l; lx >= 0; --lx)
            for (int ly = l - lx; ly >= 0; --ly) {
                s.pw.push_back(lx);
                s.pw.push_back(ly);
                s.pw.push_back(l - lx - ly);
            }
        nao += (int64_t)s.ncomp * nc;
        out.push_back(std::move(s));
    }
    return out;
}

// F_n(t) = int_0^1 u^(2n) exp(-t u^2) du for n = 0..nmax.
void boys(int nmax, double t, double* f)
{
    if (t < 30.0) {
        // series F_n(t) = e^-t sum_k (2t)^k / ((2n+1)(2n+3)...(2n+2k+1)), then down
        long double term = 1.0L / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0L * t / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < 1e-20L * sum) break;
        }
        const long double et = std::exp(-(long double)t);
        long double fn = et * sum;
        f[nmax] = (double)fn;
        for (int n = nmax - 1; n >= 0; --n) {
            fn = (2.0L * t * fn + et) / (2 * n + 1);
            f[n] = (double)fn;
        }
    } else {
        // F_0 = sqrt(pi / t) erf(sqrt t) / 2, then up (no cancellation at t >= 30)
        const double et = std::exp(-t);
        f[0] = 0.5 * std::sqrt(PI / t) * std::erf(std::sqrt(t));
        for (int n = 0; n < nmax; ++n) f[n + 1] = ((2 * n + 1) * f[n] - et) / (2 * t);
    }
}

// 1-D Hermite expansion of x_A^i x_B^j exp(-a x_A^2 - b x_B^2) / K:
// E[(i * (lb + 1) + j) * (la + lb + 1) + t].
void hermite_e(int la, int lb, double p, double xpa, double xpb, std::vector<double>& e)
{
    const int nt = la + lb + 1;
    e.assign((size_t)(la + 1) * (lb + 1) * nt, 0.0);
    auto at = [&](int i, int j, int t) -> double& { return e[((size_t)i * (lb + 1) + j) * nt + t]; };
    auto get = [&](int i, int j, int t) { return (t < 0 || t > i + j) ? 0.0 : at(i, j, t); };
    at(0, 0, 0) = 1.0;
    const double o2p = 0.5 / p;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                if (i > 0)
                    at(i, j, t) = o2p * get(i - 1, j, t - 1) + xpa * get(i - 1, j, t) + (t + 1) * get(i - 1, j, t + 1);
                else
                    at(i, j, t) = o2p * get(i, j - 1, t - 1) + xpb * get(i, j - 1, t) + (t + 1) * get(i, j - 1, t + 1);
            }
        }
}

// Hermite integrals R^0_{tuv}(alpha, X, Y, Z), t + u + v <= L, from the full
// table R^n_{tuv} (n = 0..L): r0[(t * (L + 1) + u) * (L + 1) + v].
void hermite_r(int L, double alpha, double x, double y, double z, std::vector<double>& r0)
{
    const int d = L + 1;
    std::vector<double> rn((size_t)d * d * d * d, 0.0);
    auto at = [&](int n, int t, int u, int v) -> double& { return rn[(((size_t)n * d + t) * d + u) * d + v]; };
    double f[4 * LMAX + 1];
    boys(L, alpha * (x * x + y * y + z * z), f);
    double m2a = 1.0;
    for (int n = 0; n <= L; ++n) {
        at(n, 0, 0, 0) = m2a * f[n];
        m2a *= -2.0 * alpha;
    }
    for (int total = 1; total <= L; ++total)
        for (int n = 0; n + total <= L; ++n)
            for (int t = 0; t <= total; ++t)
                for (int u = 0; t + u <= total; ++u) {
                    const int v = total - t - u;
                    double val;
                    if (t > 0)
                        val = (t > 1 ? (t - 1) * at(n + 1, t - 2, u, v) : 0.0) + x * at(n + 1, t - 1, u, v);
                    else if (u > 0)
                        val = (u > 1 ? (u - 1) * at(n + 1, t, u - 2, v) : 0.0) + y * at(n + 1, t, u - 1, v);
                    else
                        val = (v > 1 ? (v - 1) * at(n + 1, t, u, v - 2) : 0.0) + z * at(n + 1, t, u, v - 1);
                    at(n, t, u, v) = val;
                }
    r0.assign((size_t)d * d * d, 0.0);
    for (int t = 0; t <= L; ++t)
        for (int u = 0; t + u <= L; ++u)
            for (int v = 0; t + u + v <= L; ++v) r0[((size_t)t * d + u) * d + v] = at(0, t, u, v);
}

// Primitive cartesian ERI (ab|cd) of unnormalised Gaussians
// (x - A_x)^{a_x} ... exp(-alpha |r - A|^2).
double prim_eri(const int* pa, const int* pb, const int* pc, const int* pd, double ea, double eb, double ec,
                double ed, const double* A, const double* B, const double* C, const double* D)
{
    const double p = ea + eb, q = ec + ed;
    double P[3], Q[3], ab2 = 0, cd2 = 0;
    for (int k = 0; k < 3; ++k) {
        P[k] = (ea * A[k] + eb * B[k]) / p;
        Q[k] = (ec * C[k] + ed * D[k]) / q;
        ab2 += (A[k] - B[k]) * (A[k] - B[k]);
        cd2 += (C[k] - D[k]) * (C[k] - D[k]);
    }
    std::vector<double> eab[3], ecd[3];
    for (int k = 0; k < 3; ++k) {
        hermite_e(pa[k], pb[k], p, P[k] - A[k], P[k] - B[k], eab[k]);
        hermite_e(pc[k], pd[k], q, Q[k] - C[k], Q[k] - D[k], ecd[k]);
    }
    const int lab = pa[0] + pa[1] + pa[2] + pb[0] + pb[1] + pb[2];
    const int lcd = pc[0] + pc[1] + pc[2] + pd[0] + pd[1] + pd[2];
    const int L = lab + lcd, dd = L + 1;
    const double alpha = p * q / (p + q);
    std::vector<double> r;
    hermite_r(L, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], r);
    double sum = 0;
    const int tx = pa[0] + pb[0], ty = pa[1] + pb[1], tz = pa[2] + pb[2];
    const int sx = pc[0] + pd[0], sy = pc[1] + pd[1], sz = pc[2] + pd[2];
    auto E = [](const std::vector<double>& e, int la, int lb, int i, int j, int t) {
        return e[((size_t)i * (lb + 1) + j) * (la + lb + 1) + t];
    };
    for (int t = 0; t <= tx; ++t)
        for (int u = 0; u <= ty; ++u)
            for (int v = 0; v <= tz; ++v) {
                const double eb3 = E(eab[0], pa[0], pb[0], pa[0], pb[0], t) *
                                   E(eab[1], pa[1], pb[1], pa[1], pb[1], u) *
                                   E(eab[2], pa[2], pb[2], pa[2], pb[2], v);
                if (eb3 == 0) continue;
                double inner = 0;
                for (int tau = 0; tau <= sx; ++tau)
                    for (int nu = 0; nu <= sy; ++nu)
                        for (int phi = 0; phi <= sz; ++phi) {
                            const double ek = E(ecd[0], pc[0], pd[0], pc[0], pd[0], tau) *
                                              E(ecd[1], pc[1], pd[1], pc[1], pd[1], nu) *
                                              E(ecd[2], pc[2], pd[2], pc[2], pd[2], phi);
                            const double sign = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                            inner += sign * ek * r[((size_t)(t + tau) * dd + (u + nu)) * dd + (v + phi)];
                        }
                sum += eb3 * inner;
            }
    const double kab = std::exp(-ea * eb / p * ab2), kcd = std::exp(-ec * ed / q * cd2);
    return 2.0 * std::pow(PI, 2.5) / (p * q * std::sqrt(p + q)) * kab * kcd * sum;
}

// Contracted, transformed block out[((i * nB + j) * nC + k) * nD + l] over the
// AOs of shells A, B, C, D (nX = nctr * ncomp, ctr-major).
void shell_quartet(const EShell& A, const EShell& B, const EShell& C, const EShell& D, std::vector<double>& out)
{
    const EShell* s[4] = {&A, &B, &C, &D};
    const int nc[4] = {A.ncart, B.ncart, C.ncart, D.ncart};
    // cartesian block per contraction quartet, then transform
    const size_t ncq = (size_t)nc[0] * nc[1] * nc[2] * nc[3];
    std::vector<double> cartb(ncq * A.nctr * B.nctr * C.nctr * D.nctr, 0.0);
    std::vector<double> prim(ncq);
    for (int ia = 0; ia < A.nprim; ++ia)
        for (int ib = 0; ib < B.nprim; ++ib)
            for (int ic = 0; ic < C.nprim; ++ic)
                for (int id = 0; id < D.nprim; ++id) {
                    size_t q = 0;
                    for (int a = 0; a < nc[0]; ++a)
                        for (int b = 0; b < nc[1]; ++b)
                            for (int c = 0; c < nc[2]; ++c)
                                for (int d = 0; d < nc[3]; ++d)
                                    prim[q++] = prim_eri(&A.pw[3 * a], &B.pw[3 * b], &C.pw[3 * c], &D.pw[3 * d],
                                                         A.exps[ia], B.exps[ib], C.exps[ic], D.exps[id], A.c, B.c,
                                                         C.c, D.c);
                    for (int ca = 0; ca < A.nctr; ++ca)
                        for (int cb = 0; cb < B.nctr; ++cb)
                            for (int cc = 0; cc < C.nctr; ++cc)
                                for (int cd = 0; cd < D.nctr; ++cd) {
                                    const double w = A.coef[ca * A.nprim + ia] * B.coef[cb * B.nprim + ib] *
                                                     C.coef[cc * C.nprim + ic] * D.coef[cd * D.nprim + id];
                                    double* dst = &cartb[((((size_t)ca * B.nctr + cb) * C.nctr + cc) * D.nctr + cd) * ncq];
                                    for (size_t k = 0; k < ncq; ++k) dst[k] += w * prim[k];
                                }
                }
    const int n[4] = {A.nctr * A.ncomp, B.nctr * B.ncomp, C.nctr * C.ncomp, D.nctr * D.ncomp};
    out.assign((size_t)n[0] * n[1] * n[2] * n[3], 0.0);
    for (int ca = 0; ca < A.nctr; ++ca)
        for (int cb = 0; cb < B.nctr; ++cb)
            for (int cc = 0; cc < C.nctr; ++cc)
                for (int cd = 0; cd < D.nctr; ++cd) {
                    const double* blk = &cartb[((((size_t)ca * B.nctr + cb) * C.nctr + cc) * D.nctr + cd) * ncq];
                    for (int ma = 0; ma < A.ncomp; ++ma)
                        for (int mb = 0; mb < B.ncomp; ++mb)
                            for (int mc = 0; mc < C.ncomp; ++mc)
                                for (int md = 0; md < D.ncomp; ++md) {
                                    double v = 0;
                                    size_t q = 0;
                                    for (int a = 0; a < nc[0]; ++a)
                                        for (int b = 0; b < nc[1]; ++b)
                                            for (int c = 0; c < nc[2]; ++c)
                                                for (int d = 0; d < nc[3]; ++d, ++q)
                                                    v += s[0]->t[ma * nc[0] + a] * s[1]->t[mb * nc[1] + b] *
                                                         s[2]->t[mc * nc[2] + c] * s[3]->t[md * nc[3] + d] * blk[q];
                                    const int i = ca * A.ncomp + ma, j = cb * B.ncomp + mb;
                                    const int k = cc * C.ncomp + mc, l = cd * D.ncomp + md;
                                    out[(((size_t)i * n[1] + j) * n[2] + k) * n[3] + l] = v;
                                }
                }
}

size_t pair_index(size_t a, size_t b) { return a * (a + 1) / 2 + b; }

std::vector<int> ao_to_shell(const std::vector<EShell>& sh, int64_t nao)
{
    std::vector<int> m((size_t)nao);
    for (size_t s = 0; s < sh.size(); ++s)
        for (int a = 0; a < sh[s].nctr * sh[s].ncomp; ++a) m[sh[s].ao_off + a] = (int)s;
    return m;
}

// Packed column (xy|mu nu) for every pair x >= y (cholesky_eri_greedy.cpp:121-140).
void pivot_column(const std::vector<EShell>& sh, const std::vector<int>& owner, size_t mu, size_t nu,
                  std::vector<double>& column)
{
    const EShell& a = sh[owner[mu]];
    const EShell& b = sh[owner[nu]];
    const size_t ai = mu - a.ao_off, bi = nu - b.ao_off;
    const int na = a.nctr * a.ncomp, nb = b.nctr * b.ncomp;
    std::vector<std::pair<int, int>> pairs;
    for (size_t s = 0; s < sh.size(); ++s)
        for (size_t t = 0; t <= s; ++t) pairs.emplace_back((int)s, (int)t);
#pragma omp parallel for schedule(dynamic)
    for (size_t task = 0; task < pairs.size(); ++task) {
        const EShell& c = sh[pairs[task].first];
        const EShell& d = sh[pairs[task].second];
        std::vector<double> blk;
        shell_quartet(c, d, a, b, blk);
        const int nc = c.nctr * c.ncomp, nd = d.nctr * d.ncomp;
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nd; ++j) {
                const size_t x = c.ao_off + i, y = d.ao_off + j;
                if (x < y) continue;
                column[pair_index(x, y)] = blk[(((size_t)i * nd + j) * na + ai) * nb + bi];
            }
    }
}

struct Decomposition {
    size_t npair = 0;
    std::vector<double> factors;   // [rank][npair]
    std::vector<int64_t> pivots;
    std::vector<double> residual;
};

Decomposition cholesky(const std::vector<EShell>& sh, int64_t nao, double tolerance, int64_t max_rank,
                       int algorithm, uint64_t seed)
{
    if (algorithm != 0 && algorithm != 1) throw std::invalid_argument("unknown Cholesky algorithm");
    if (!std::isfinite(tolerance) || tolerance < 0)
        throw std::invalid_argument("Cholesky tolerance must be finite and nonnegative");
    const size_t npair = (size_t)nao * (nao + 1) / 2;
    if (npair == 0) throw std::invalid_argument("empty AO basis");
    if (max_rank < 0 || (size_t)max_rank > npair) throw std::invalid_argument("max_rank exceeds AO-pair dimension");
    const auto owner = ao_to_shell(sh, nao);
    Decomposition r;
    r.npair = npair;
    r.residual.assign(npair, 0.0);
    // diagonal (xy|xy) from the (st|st) quartets (cholesky_eri_greedy.cpp:80-96)
    for (size_t s = 0; s < sh.size(); ++s)
        for (size_t t = 0; t <= s; ++t) {
            std::vector<double> blk;
            shell_quartet(sh[s], sh[t], sh[s], sh[t], blk);
            const int na = sh[s].nctr * sh[s].ncomp, nb = sh[t].nctr * sh[t].ncomp;
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const size_t x = sh[s].ao_off + i, y = sh[t].ao_off + j;
                    if (x < y) continue;
                    const double v = blk[(((size_t)i * nb + j) * na + i) * nb + j];
                    if (!std::isfinite(v) || v < -1e-12) throw std::runtime_error("invalid ERI diagonal");
                    r.residual[pair_index(x, y)] = std::max(0.0, v);
                }
        }
    const auto first = std::max_element(r.residual.begin(), r.residual.end());  // first of equal maxima
    const double cutoff = tolerance * *first;
    const size_t cap = max_rank ? (size_t)max_rank : npair;
    std::mt19937_64 gen(seed);
    double next_res = *first;
    size_t next_idx = (size_t)(first - r.residual.begin());
    std::vector<double> column(npair);
    while (r.pivots.size() < cap) {
        if (next_res <= cutoff || next_res <= 0.0) break;
        size_t pivot = next_idx;
        double pres = next_res;
        if (algorithm == 1) {  // uniform among residuals above the cutoff (cholesky_pivot.hpp:16-36)
            size_t eligible = 0;
            for (double v : r.residual)
                if (v > cutoff && v > 0.0) ++eligible;
            std::uniform_int_distribution<size_t> pick(0, eligible - 1);
            size_t ordinal = pick(gen);
            for (size_t p = 0; p < npair; ++p)
                if (r.residual[p] > cutoff && r.residual[p] > 0.0 && ordinal-- == 0) {
                    pivot = p;
                    pres = r.residual[p];
                    break;
                }
        }
        size_t lo = 0, hi = (size_t)nao;
        while (lo + 1 < hi) {
            const size_t mid = lo + (hi - lo) / 2;
            if (pair_index(mid, 0) <= pivot) lo = mid; else hi = mid;
        }
        const size_t mu = lo, nu = pivot - pair_index(mu, 0);
        std::fill(column.begin(), column.end(), 0.0);
        pivot_column(sh, owner, mu, nu, column);
        // correct_cholesky_column: k in order, then the next maximum (first index wins ties)
        const size_t rank = r.pivots.size();
        std::vector<double> pf(rank);
        for (size_t k = 0; k < rank; ++k) pf[k] = r.factors[k * npair + pivot];
        r.factors.resize((rank + 1) * npair);
        const double denom = std::sqrt(pres);
        next_res = -1.0;
        next_idx = npair;
        for (size_t p = 0; p < npair; ++p) {
            double v = column[p];
            for (size_t k = 0; k < rank; ++k) v -= r.factors[k * npair + p] * pf[k];
            v /= denom;
            r.factors[rank * npair + p] = v;
            if (!std::isfinite(v)) throw std::runtime_error("non-finite Cholesky vector");
            r.residual[p] = p == pivot ? 0.0 : std::max(0.0, r.residual[p] - v * v);
            if (r.residual[p] > next_res) {
                next_res = r.residual[p];
                next_idx = p;
            }
        }
        r.pivots.push_back((int64_t)pivot);
    }
    return r;
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

const char* oracle_eri_last_error() { return g_error.c_str(); }

// Primitive cartesian ERI: pow[12] = (a_x, a_y, a_z, b_x, ..., d_z),
// exps[4], xyz[12] (A, B, C, D).
int oracle_eri_prim(const int* pow, const double* exps, const double* xyz, double* out)
{
    return guarded([&] {
        for (int i = 0; i < 12; ++i)
            if (pow[i] < 0 || pow[i] > LMAX) throw std::invalid_argument("power out of range");
        *out = prim_eri(pow, pow + 3, pow + 6, pow + 9, exps[0], exps[1], exps[2], exps[3], xyz, xyz + 3, xyz + 6,
                        xyz + 9);
    });
}

// Boys function F_0..F_nmax at t.
int oracle_boys(int nmax, double t, double* out)
{
    return guarded([&] {
        if (nmax < 0 || nmax > 4 * LMAX || !(t >= 0)) throw std::invalid_argument("boys arguments");
        boys(nmax, t, out);
    });
}

// Packed ERI matrix out[pair(mu,nu) * npair + pair(la,si)] over mu >= nu, la >= si.
int oracle_eri_packed(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env, int64_t nenv,
                      int cart, double* out)
{
    return guarded([&] {
        int64_t nao = 0;
        const auto sh = shells_of(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        const size_t npair = (size_t)nao * (nao + 1) / 2;
        std::vector<std::pair<int, int>> pairs;
        for (size_t s = 0; s < sh.size(); ++s)
            for (size_t t = 0; t <= s; ++t) pairs.emplace_back((int)s, (int)t);
#pragma omp parallel for schedule(dynamic)
        for (size_t i = 0; i < pairs.size(); ++i)
            for (size_t j = 0; j < pairs.size(); ++j) {
                const EShell& a = sh[pairs[i].first];
                const EShell& b = sh[pairs[i].second];
                const EShell& c = sh[pairs[j].first];
                const EShell& d = sh[pairs[j].second];
                std::vector<double> blk;
                shell_quartet(a, b, c, d, blk);
                const int na = a.nctr * a.ncomp, nb = b.nctr * b.ncomp, nc = c.nctr * c.ncomp, nd = d.nctr * d.ncomp;
                for (int p = 0; p < na; ++p)
                    for (int q = 0; q < nb; ++q) {
                        const size_t x = a.ao_off + p, y = b.ao_off + q;
                        if (x < y) continue;
                        for (int r = 0; r < nc; ++r)
                            for (int s2 = 0; s2 < nd; ++s2) {
                                const size_t z = c.ao_off + r, w = d.ao_off + s2;
                                if (z < w) continue;
                                out[pair_index(x, y) * npair + pair_index(z, w)] =
                                    blk[(((size_t)p * nb + q) * nc + r) * nd + s2];
                            }
                    }
            }
    });
}

// cholesky_eri: factors[capacity][npair] (first *rank rows), pivots, residual[npair].
int oracle_cholesky_eri(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env, int64_t nenv,
                        int cart, double tolerance, int64_t max_rank, int algorithm, uint64_t seed, int64_t capacity,
                        double* factors, int64_t* pivots, double* residual, int64_t* rank)
{
    return guarded([&] {
        int64_t nao = 0;
        const auto sh = shells_of(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        const auto r = cholesky(sh, nao, tolerance, max_rank, algorithm, seed);
        if ((int64_t)r.pivots.size() > capacity) throw std::overflow_error("Cholesky rank exceeds capacity");
        std::copy(r.factors.begin(), r.factors.end(), factors);
        std::copy(r.pivots.begin(), r.pivots.end(), pivots);
        std::copy(r.residual.begin(), r.residual.end(), residual);
        *rank = (int64_t)r.pivots.size();
    });
}

// coulomb_metric: out[nvec][nvec] = P P^T with P = packed(transform) F^T
// (coulomb_metric.cpp:20-98; transform is row-major [nvec][nao * nao]).
int oracle_coulomb_metric(const int32_t* atm, int natm, const int32_t* bas, int nbas, const double* env,
                          int64_t nenv, int cart, const double* transform, int64_t nvec, int algorithm,
                          double tolerance, int64_t max_rank, int64_t batch_size, uint64_t seed, double* out)
{
    return guarded([&] {
        int64_t nao = 0;
        const auto sh = shells_of(atm, natm, bas, nbas, env, nenv, cart != 0, nao);
        if (batch_size <= 0) throw std::invalid_argument("batch_size must be positive");
        for (int64_t i = 0; i < nvec * nao * nao; ++i)
            if (!std::isfinite(transform[i])) throw std::invalid_argument("transform must be finite");
        const auto r = cholesky(sh, nao, tolerance, max_rank, algorithm, seed);
        const size_t npair = r.npair, rank = r.pivots.size();
        std::fill(out, out + nvec * nvec, 0.0);
        if (rank == 0 || nvec == 0) return;
        std::vector<double> proj((size_t)nvec * rank, 0.0), packed(npair);
        for (int64_t v = 0; v < nvec; ++v) {
            for (int64_t mu = 0; mu < nao; ++mu)
                for (int64_t nu = 0; nu <= mu; ++nu)
                    packed[pair_index(mu, nu)] = transform[(v * nao + mu) * nao + nu] +
                                                 (mu == nu ? 0.0 : transform[(v * nao + nu) * nao + mu]);
            for (size_t k = 0; k < rank; ++k) {
                double s = 0;
                for (size_t p = 0; p < npair; ++p) s += packed[p] * r.factors[k * npair + p];
                proj[v * rank + k] = s;
            }
        }
        for (int64_t a = 0; a < nvec; ++a)
            for (int64_t b = 0; b < nvec; ++b) {
                double s = 0;
                for (size_t k = 0; k < rank; ++k) s += proj[a * rank + k] * proj[b * rank + k];
                out[a * nvec + b] = s;
            }
    });
}

}  // extern "C"
