// Lebedev-Laikov angular quadrature rules on the unit sphere: the angular
// part of PySCF's dft.gen_grid.Grids, which the reference takes its grid
// from (python/isdf.py:488-496).
#pragma once

#include <vector>

namespace mpisdf {

// One orbit of the octahedral group O_h: kind 1 (1,0,0) 6 points, 2
// (0,1,1)/sqrt2 12, 3 (1,1,1)/sqrt3 8, 4 (l,l,m) 24, 5 (p,q,0) 24,
// 6 (r,s,t) 48; a, b are the free coordinates (l; p; r, s), w the weight of
// each of its points (the weights of a rule sum to 1).
struct LebedevOrbit {
    int kind;
    double a, b, w;
};

struct LebedevRule {
    int npoints, degree, norbits;
    const LebedevOrbit* orbits;
};

// The tabulated rule with npoints points, or nullptr.
const LebedevRule* lebedev_rule(int npoints);
// Points (unit vectors, xyz[3 n]) and weights (summing to 1) of a rule, orbit
// by orbit.
void lebedev_points(const LebedevRule& rule, std::vector<double>& xyz, std::vector<double>& w);
// Sizes of the tabulated rules, ascending.
std::vector<int> lebedev_sizes();

}  // namespace mpisdf
