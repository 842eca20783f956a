// Molecular integration grid (host, multithreaded): the input the reference
// takes from PySCF's dft.gen_grid.Grids (python/isdf.py:488-496), built here
// without PySCF, with PySCF's defaults:
//
// * Radial: Treutler-Ahlrichs M4 mapping of Gauss-Chebyshev (2nd kind)
//   nodes, r = xi/ln2 (1+x)^0.6 ln(2/(1-x)), xi_H = 0.8, xi_O = 0.9; the
//   number of shells per level and period (RAD_GRIDS).
// * Angular: Lebedev-Laikov rules (lebedev.cpp) of the order PySCF assigns
//   per level and period (ANG_ORDER), pruned NWChem-style: five regions on
//   r / R_Bragg take the rules [1, 3, i-1, i, i-1] of the Lebedev size list
//   from 38 points up, i being the full rule's position ([1, 2, 2, 2, 1] for
//   the 50-point rule).
// * Partition: Becke fuzzy cells (three smoothing iterations) with Treutler's
//   sqrt(R_Bragg) size adjustment, a_ij clipped to [-1/2, 1/2].
// * Order: atom, radial shell, Lebedev orbit -- consecutive points are
//   spatially close, which the AO kernel's tile culling exploits.
// * The partition runs on the GPU when one is present (one thread per point)
//   and on host threads otherwise; both evaluate the same function with no
//   FMA contraction, so the weights are identical.
#include "grid.hpp"

#include "common.hpp"
#include "lebedev.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace mpisdf {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kBohrPerAngstrom = 1.0 / 0.52917721092;

double bragg_radius_bohr(int z)
{
    // Bragg-Slater radii (Angstrom); hydrogen 0.35 as in PySCF's table.
    static const double table[] = {0.0, 0.35, 1.40, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 1.50};
    if (z < 1 || z > 10) throw std::invalid_argument("grid generator supports Z = 1..10");
    return table[z] * kBohrPerAngstrom;
}

double treutler_xi(int z)
{
    static const double table[] = {0.0, 0.8, 0.9, 1.8, 1.4, 1.3, 1.1, 0.9, 0.9, 0.9, 0.9};
    return table[z];
}

int period(int z) { return z <= 2 ? 0 : 1; }

int radial_points(int z, int level)
{
    static const int table[2][10] = {{10, 30, 40, 50, 60, 70, 80, 90, 100, 200},
                                     {15, 40, 60, 75, 90, 105, 120, 135, 150, 200}};
    return table[period(z)][std::clamp(level, 0, 9)];
}

// Lebedev sizes from 38 points up (PySCF's LEBEDEV_NGRID[4:]); the pruning
// regions index into this list.
constexpr int kLebedevFrom38[] = {38,   50,   74,   86,   110,  146,  170,  194,  230,  266,
                                  302,  350,  434,  590,  770,  974,  1202, 1454, 1730, 2030,
                                  2354, 2702, 3074, 3470, 3890, 4334, 4802, 5294, 5810};

// Angular rule size of the full (unpruned) shells: ANG_ORDER[level][period]
// mapped through the Lebedev order -> size table.
int angular_size(int z, int level)
{
    static const int sizes[2][10] = {{50, 110, 194, 302, 434, 590, 770, 974, 1202, 1454},
                                     {86, 194, 302, 302, 590, 770, 974, 1202, 1202, 1454}};
    return sizes[period(z)][std::clamp(level, 0, 9)];
}

// Rule sizes of the five NWChem pruning regions for a full size n.
void pruned_sizes(int n, int out[5])
{
    if (n < 50) {
        std::fill(out, out + 5, n);
        return;
    }
    int idx[5] = {1, 2, 2, 2, 1};
    if (n != 50) {
        int i = 0;
        while (kLebedevFrom38[i] != n) ++i;
        idx[0] = 1, idx[1] = 3, idx[2] = i - 1, idx[3] = i, idx[4] = i - 1;
    }
    for (int r = 0; r < 5; ++r) out[r] = kLebedevFrom38[idx[r]];
}

struct Angular {
    std::vector<double> xyz, w;  // unit vectors; weights sum to 4 pi
};

Angular angular_rule(int n, int level)
{
    const LebedevRule* rule = lebedev_rule(n);
    if (!rule)
        throw std::invalid_argument("grid level " + std::to_string(level) + " needs the " + std::to_string(n) +
                                    "-point Lebedev rule, which is not tabulated");
    Angular a;
    lebedev_points(*rule, a.xyz, a.w);
    for (double& x : a.w) x *= 4 * kPi;
    return a;
}

struct AtomGrid {
    std::vector<double> xyz, w;  // relative to the atom, quadrature weight (no partition)
};

AtomGrid atom_grid(int z, int level)
{
    const int nr = radial_points(z, level);
    const double xi = treutler_xi(z);
    const double rb = bragg_radius_bohr(z);
    // NWChem-style pruning regions (r / R_Bragg) and Lebedev sizes per region.
    static const double alphas[2][4] = {{0.25, 0.5, 1.0, 4.5}, {0.1667, 0.5, 0.9, 3.5}};
    int sizes[5];
    pruned_sizes(angular_size(z, level), sizes);
    AtomGrid g;
    std::vector<Angular> rules(5);
    for (int r = 0; r < 5; ++r) rules[r] = angular_rule(sizes[r], level);
    for (int i = 1; i <= nr; ++i) {
        const double th = i * kPi / (nr + 1);
        const double x = std::cos(th);
        const double lx = std::log(2.0 / (1 - x));
        const double px = std::pow(1 + x, 0.6);
        const double r = xi / std::log(2.0) * px * lx;
        const double dr = xi / std::log(2.0) * (0.6 * std::pow(1 + x, -0.4) * lx + px / (1 - x));
        const double wr = kPi / (nr + 1) * std::sin(th) * dr * r * r;
        int place = 0;
        for (int k = 0; k < 4; ++k)
            if (r / rb > alphas[period(z)][k]) ++place;
        const Angular& a = rules[place];
        for (size_t j = 0; j < a.w.size(); ++j) {
            for (int c = 0; c < 3; ++c) g.xyz.push_back(r * a.xyz[3 * j + c]);
            g.w.push_back(wr * a.w[j]);
        }
    }
    return g;
}

template <class F>
void parallel_for(int64_t n, F&& f)
{
    const int nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 64u));
    std::vector<std::thread> pool;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&, b, e] { for (int64_t i = b; i < e; ++i) f(i); });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

namespace {

struct Layout {
    std::vector<AtomGrid> per_element;
    std::vector<int64_t> offset;  // first point of each atom
    std::vector<double> rpair, apair;
};

Layout layout(const int* charges, const double* coords, int64_t natm, int level)
{
    if (natm <= 0) throw std::invalid_argument("grid needs at least one atom");
    if (level < 0 || level > 9) throw std::invalid_argument("grid level must be 0..9");
    Layout L;
    L.per_element.resize(11);
    std::vector<double> sqrt_r(natm);
    L.rpair.resize((size_t)(natm * natm));
    L.apair.resize((size_t)(natm * natm));
    for (int64_t i = 0; i < natm; ++i) {
        const int z = charges[i];
        bragg_radius_bohr(z);  // validates Z before indexing
        if (L.per_element[z].w.empty()) L.per_element[z] = atom_grid(z, level);
        sqrt_r[i] = std::sqrt(bragg_radius_bohr(z));
    }
    for (int64_t i = 0; i < natm; ++i)
        for (int64_t j = 0; j < natm; ++j) {
            const double dx = coords[3 * i] - coords[3 * j], dy = coords[3 * i + 1] - coords[3 * j + 1],
                         dz = coords[3 * i + 2] - coords[3 * j + 2];
            L.rpair[i * natm + j] = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double rr = sqrt_r[i] / sqrt_r[j];
            L.apair[i * natm + j] = std::clamp(0.25 * (1 / rr - rr), -0.5, 0.5);
        }
    for (int64_t i = 0; i < natm; ++i)
        for (int64_t j = 0; j < i; ++j)
            if (L.rpair[i * natm + j] < 1e-8) throw std::invalid_argument("atoms must not coincide");
    L.offset.assign(natm + 1, 0);
    for (int64_t i = 0; i < natm; ++i) L.offset[i + 1] = L.offset[i] + (int64_t)L.per_element[charges[i]].w.size();
    return L;
}

// The Becke partition of one point, written once for the host loop and the
// device kernel so both round every operation alike (no contraction into
// FMAs): w = wq * P_atom / sum_i P_i over the cells that can contribute.
#pragma clang fp contract(off)
__host__ __device__ inline double becke_s_point(double mu)
{
    double g = mu;
    for (int i = 0; i < 3; ++i) g = 1.5 * g - 0.5 * g * g * g;
    return 0.5 * (1 - g);
}

__host__ __device__ inline double atom_distance(const double* coords, int64_t i, double px, double py, double pz)
{
    const double dx = px - coords[3 * i], dy = py - coords[3 * i + 1], dz = pz - coords[3 * i + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

__host__ __device__ inline double becke_point(const double* coords, const double* rpair, const double* apair,
                                              int64_t natm, int64_t atom, double px, double py, double pz, double wq)
{
    int64_t nearest = 0;
    double dn = atom_distance(coords, 0, px, py, pz);
    for (int64_t i = 1; i < natm; ++i) {
        const double di = atom_distance(coords, i, px, py, pz);
        if (di < dn) { nearest = i; dn = di; }
    }
    double sum = 0, own = 0;
    for (int64_t i = 0; i < natm; ++i) {
        const double di = atom_distance(coords, i, px, py, pz);
        if (i != nearest && i != atom) {
            // Cells whose factor against the nearest atom is below ~1e-15
            // (adjusted mu > 0.99) cannot contribute.
            double mu = (di - dn) / rpair[i * natm + nearest];
            mu += apair[i * natm + nearest] * (1 - mu * mu);
            if (!(mu < 0.99)) continue;
        }
        double prod = 1;
        for (int64_t j = 0; j < natm && prod > 0; ++j) {
            if (j == i) continue;
            double mu = (di - atom_distance(coords, j, px, py, pz)) / rpair[i * natm + j];
            mu += apair[i * natm + j] * (1 - mu * mu);
            prod *= becke_s_point(mu);
        }
        sum += prod;
        if (i == atom) own = prod;
    }
    return sum > 0 ? wq * own / sum : 0.0;
}

__global__ void becke_partition_kernel(const double* __restrict__ xyz, const double* __restrict__ wq,
                                       const int32_t* __restrict__ atom_of, const double* __restrict__ coords,
                                       const double* __restrict__ rpair, const double* __restrict__ apair,
                                       int64_t natm, int64_t n, double* __restrict__ w)
{
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        w[p] = becke_point(coords, rpair, apair, natm, atom_of[p], xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2], wq[p]);
}
#pragma clang fp contract(on)

bool have_device()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

}  // namespace

int64_t becke_grid_size(const int* charges, int64_t natm, int level)
{
    if (natm <= 0) throw std::invalid_argument("grid needs at least one atom");
    if (level < 0 || level > 9) throw std::invalid_argument("grid level must be 0..9");
    std::vector<int64_t> per(11, -1);
    int64_t total = 0;
    for (int64_t i = 0; i < natm; ++i) {
        const int z = charges[i];
        bragg_radius_bohr(z);
        if (per[z] < 0) per[z] = (int64_t)atom_grid(z, level).w.size();
        total += per[z];
    }
    return total;
}

std::pair<std::vector<double>, std::vector<double>> becke_grid(const int* charges, const double* coords, int64_t natm,
                                                               int level, bool device)
{
    const Layout L = layout(charges, coords, natm, level);
    const int64_t total = L.offset[natm];
    std::vector<double> xyz((size_t)(3 * total)), wq((size_t)total), w((size_t)total);
    std::vector<int32_t> atom_of((size_t)total);
    for (int64_t atom = 0; atom < natm; ++atom) {
        const AtomGrid& g = L.per_element[charges[atom]];
        for (int64_t p = 0; p < (int64_t)g.w.size(); ++p) {
            const int64_t out = L.offset[atom] + p;
            for (int c = 0; c < 3; ++c) xyz[3 * out + c] = coords[3 * atom + c] + g.xyz[3 * p + c];
            wq[out] = g.w[p];
            atom_of[out] = (int32_t)atom;
        }
    }
    if (device && have_device()) {
        // One thread per point; the partition is O(n natm^2) in the worst
        // case, seconds on the host at 2e6 points and 192 atoms.
        DeviceBuffer<double> dx, dwq, dc, drp, dap, dw(total);
        DeviceBuffer<int32_t> da;
        dx.upload(xyz.data(), xyz.size(), nullptr);
        dwq.upload(wq.data(), wq.size(), nullptr);
        dc.upload(coords, 3 * natm, nullptr);
        drp.upload(L.rpair.data(), L.rpair.size(), nullptr);
        dap.upload(L.apair.data(), L.apair.size(), nullptr);
        da.upload(atom_of.data(), atom_of.size(), nullptr);
        const int64_t blocks = std::min<int64_t>((total + 255) / 256, 1 << 20);
        hipLaunchKernelGGL(becke_partition_kernel, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(256), 0, nullptr,
                           dx.ptr, dwq.ptr, da.ptr, dc.ptr, drp.ptr, dap.ptr, natm, total, dw.ptr);
        MPISDF_HIP(hipGetLastError());
        MPISDF_HIP(hipMemcpy(w.data(), dw.ptr, total * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        parallel_for(total, [&](int64_t p) {
            w[p] = becke_point(coords, L.rpair.data(), L.apair.data(), natm, atom_of[p], xyz[3 * p], xyz[3 * p + 1],
                               xyz[3 * p + 2], wq[p]);
        });
    }
    return {std::move(xyz), std::move(w)};
}

}  // namespace mpisdf
