// Lebedev-Laikov rules (V. I. Lebedev and D. N. Laikov, Doklady Mathematics
// 59 (1999) 477): octahedrally symmetric point sets whose orbit coordinates
// and weights solve the moment equations for every O_h-invariant spherical
// harmonic up to the rule's degree.  The table is data written by
// lebedev_solve.py (beside this file), which solves those equations; the
// points are generated from the orbit generators here.
#include "lebedev.hpp"

#include <cmath>

namespace mpisdf {
namespace {

#include "lebedev_rules.inc"

void push(std::vector<double>& xyz, std::vector<double>& w, double x, double y, double z, double wt)
{
    xyz.push_back(x);
    xyz.push_back(y);
    xyz.push_back(z);
    w.push_back(wt);
}

// All sign combinations of (x, y, z), skipping sign flips of zero coordinates.
void signs(std::vector<double>& xyz, std::vector<double>& w, double x, double y, double z, double wt)
{
    for (int sx = 0; sx < (x != 0 ? 2 : 1); ++sx)
        for (int sy = 0; sy < (y != 0 ? 2 : 1); ++sy)
            for (int sz = 0; sz < (z != 0 ? 2 : 1); ++sz)
                push(xyz, w, sx ? -x : x, sy ? -y : y, sz ? -z : z, wt);
}

}  // namespace

const LebedevRule* lebedev_rule(int npoints)
{
    for (const auto& r : kLebedevRules)
        if (r.npoints == npoints) return &r;
    return nullptr;
}

std::vector<int> lebedev_sizes()
{
    std::vector<int> out;
    for (const auto& r : kLebedevRules) out.push_back(r.npoints);
    return out;
}

void lebedev_points(const LebedevRule& rule, std::vector<double>& xyz, std::vector<double>& w)
{
    xyz.clear();
    w.clear();
    for (int o = 0; o < rule.norbits; ++o) {
        const LebedevOrbit& g = rule.orbits[o];
        switch (g.kind) {
        case 1:
            signs(xyz, w, 1, 0, 0, g.w);
            signs(xyz, w, 0, 1, 0, g.w);
            signs(xyz, w, 0, 0, 1, g.w);
            break;
        case 2: {
            const double a = std::sqrt(0.5);
            signs(xyz, w, 0, a, a, g.w);
            signs(xyz, w, a, 0, a, g.w);
            signs(xyz, w, a, a, 0, g.w);
            break;
        }
        case 3: {
            const double a = std::sqrt(1.0 / 3.0);
            signs(xyz, w, a, a, a, g.w);
            break;
        }
        case 4: {
            const double l = g.a, m = std::sqrt(1 - 2 * l * l);
            signs(xyz, w, l, l, m, g.w);
            signs(xyz, w, l, m, l, g.w);
            signs(xyz, w, m, l, l, g.w);
            break;
        }
        case 5: {
            const double p = g.a, q = std::sqrt(1 - p * p);
            signs(xyz, w, p, q, 0, g.w);
            signs(xyz, w, q, p, 0, g.w);
            signs(xyz, w, p, 0, q, g.w);
            signs(xyz, w, q, 0, p, g.w);
            signs(xyz, w, 0, p, q, g.w);
            signs(xyz, w, 0, q, p, g.w);
            break;
        }
        default: {
            const double r = g.a, s = g.b, t = std::sqrt(1 - r * r - s * s);
            signs(xyz, w, r, s, t, g.w);
            signs(xyz, w, r, t, s, g.w);
            signs(xyz, w, s, r, t, g.w);
            signs(xyz, w, s, t, r, g.w);
            signs(xyz, w, t, r, s, g.w);
            signs(xyz, w, t, s, r, g.w);
            break;
        }
        }
    }
}

}  // namespace mpisdf
