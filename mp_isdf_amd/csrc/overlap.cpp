// Analytic AO overlap matrix of a basis context: the counterpart of PySCF's
// mol.intor("int1e_ovlp") that the reference's overlap check compares with
// (python/isdf.py:665-671).
//
// Cartesian primitive overlaps come from the Obara-Saika recursion in each
// direction,
//   s(0,0)     = sqrt(pi/p) exp(-a b / p (A - B)^2),
//   s(i+1, j)  = X_PA s(i,j) + (i s(i-1,j) + j s(i,j-1)) / (2p),
//   s(i, j+1)  = X_PB s(i,j) + (i s(i-1,j) + j s(i,j-1)) / (2p),
// with p = a + b and P = (a A + b B) / p.  They are contracted with the
// coefficients of the context's env and transformed by the same per-l
// cartesian-to-AO tables the AO kernel applies, so S is the overlap of
// exactly the functions eval_ao evaluates.  Host code: O(nshell^2 nprim^2
// ncart^2) work, well under a second for 1536 AOs.
#include "context.hpp"

#include <cmath>
#include <vector>

namespace mpisdf {
namespace {

constexpr double kPi = 3.14159265358979323846;

// s[i * (lb + 1) + j] for i <= la, j <= lb.
void os_1d(double a, double b, double A, double B, int la, int lb, double* s)
{
    const double p = a + b, P = (a * A + b * B) / p;
    const double xpa = P - A, xpb = P - B, h = 0.5 / p;
    const int ld = lb + 1;
    s[0] = std::sqrt(kPi / p) * std::exp(-a * b / p * (A - B) * (A - B));
    for (int i = 0; i < la; ++i)
        s[(i + 1) * ld] = xpa * s[i * ld] + (i > 0 ? h * i * s[(i - 1) * ld] : 0.0);
    for (int j = 0; j < lb; ++j)
        for (int i = 0; i <= la; ++i)
            s[i * ld + j + 1] = xpb * s[i * ld + j] + h * ((i > 0 ? i * s[(i - 1) * ld + j] : 0.0) +
                                                           (j > 0 ? j * s[i * ld + j - 1] : 0.0));
}

// Cartesian exponents of component c of angular momentum l in the AO
// kernel's order (lx from l down, then ly from l - lx down).
void cart_exponents(int l, std::vector<int>& ex, std::vector<int>& ey, std::vector<int>& ez)
{
    ex.clear();
    ey.clear();
    ez.clear();
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            ex.push_back(lx);
            ey.push_back(ly);
            ez.push_back(l - lx - ly);
        }
}

}  // namespace

void overlap_matrix(const Context& ctx, double* S)
{
    const int64_t nao = ctx.nao;
    std::vector<int> axe, aye, aze, bxe, bye, bze;
    std::vector<double> sx, sy, sz, cart, half;
    for (int ia = 0; ia < ctx.nshells; ++ia) {
        const DevShell& A = ctx.shells[ia];
        const int nca = (A.l + 1) * (A.l + 2) / 2;
        cart_exponents(A.l, axe, aye, aze);
        const double* ta = ctx.tabs.data() + A.tab;
        for (int ib = 0; ib < ctx.nshells; ++ib) {
            const DevShell& B = ctx.shells[ib];
            const int ncb = (B.l + 1) * (B.l + 2) / 2;
            cart_exponents(B.l, bxe, bye, bze);
            const double* tb = ctx.tabs.data() + B.tab;
            const int ld = B.l + 1;
            sx.resize((A.l + 1) * ld);
            sy.resize((A.l + 1) * ld);
            sz.resize((A.l + 1) * ld);
            // cart[(ca * nctr_b + cb_ctr) ...]: contracted cartesian overlaps
            // per contraction pair, [ctr_a][ctr_b][nca][ncb].
            cart.assign((size_t)A.nctr * B.nctr * nca * ncb, 0.0);
            for (int pa = 0; pa < A.nprim; ++pa) {
                const double a = ctx.env[A.exp_off + pa];
                for (int pb = 0; pb < B.nprim; ++pb) {
                    const double b = ctx.env[B.exp_off + pb];
                    os_1d(a, b, A.x, B.x, A.l, B.l, sx.data());
                    os_1d(a, b, A.y, B.y, A.l, B.l, sy.data());
                    os_1d(a, b, A.z, B.z, A.l, B.l, sz.data());
                    for (int ca = 0; ca < A.nctr; ++ca) {
                        const double wa = ctx.env[A.coeff_off + ca * A.nprim + pa];
                        for (int cb = 0; cb < B.nctr; ++cb) {
                            const double wab = wa * ctx.env[B.coeff_off + cb * B.nprim + pb];
                            double* blk = cart.data() + ((size_t)ca * B.nctr + cb) * nca * ncb;
                            for (int u = 0; u < nca; ++u)
                                for (int v = 0; v < ncb; ++v)
                                    blk[u * ncb + v] += wab * sx[axe[u] * ld + bxe[v]] * sy[aye[u] * ld + bye[v]] *
                                                        sz[aze[u] * ld + bze[v]];
                        }
                    }
                }
            }
            // S block = T_a cart T_b^T per contraction pair.
            half.resize((size_t)A.ncomp * ncb);
            for (int ca = 0; ca < A.nctr; ++ca)
                for (int cb = 0; cb < B.nctr; ++cb) {
                    const double* blk = cart.data() + ((size_t)ca * B.nctr + cb) * nca * ncb;
                    for (int s = 0; s < A.ncomp; ++s)
                        for (int v = 0; v < ncb; ++v) {
                            double acc = 0.0;
                            for (int u = 0; u < nca; ++u) acc += ta[s * nca + u] * blk[u * ncb + v];
                            half[s * ncb + v] = acc;
                        }
                    for (int s = 0; s < A.ncomp; ++s)
                        for (int t = 0; t < B.ncomp; ++t) {
                            double acc = 0.0;
                            for (int v = 0; v < ncb; ++v) acc += half[s * ncb + v] * tb[t * ncb + v];
                            S[(int64_t)(A.ao_off + ca * A.ncomp + s) * nao + B.ao_off + cb * B.ncomp + t] = acc;
                        }
                }
        }
    }
}

}  // namespace mpisdf
