"""Solve for the Lebedev-Laikov angular rules and write lebedev_rules.inc.

A Lebedev rule of degree L is a set of orbits of the octahedral group O_h
(V. I. Lebedev and D. N. Laikov, Doklady Mathematics 59 (1999) 477):

    a1 (1,0,0) 6 points, a2 (0,1,1)/sqrt2 12, a3 (1,1,1)/sqrt3 8,
    b (l,l,m) 24, c (p,q,0) 24, d (r,s,t) 48,

with one weight per orbit, such that every spherical harmonic of degree <= L
is integrated exactly.  Odd degrees vanish by symmetry and so do harmonics
that are not O_h-invariant, so the conditions are the moments of an
orthonormal basis of the invariant harmonics of even degree < L, one equation
per basis function.  For each rule the orbit structure below has exactly as
many unknowns (weights and free coordinates) as equations.

The weights enter linearly and are eliminated by least squares (variable
projection); the free coordinates are found by Levenberg-Marquardt from
spread starting points, raising the degree of the moment conditions in
steps of four (a homotopy that keeps the iterates near well-distributed
point sets).  The first converged solution with all weights positive is
kept; the 74- and 266-point rules have no positive solution with their
structure and keep the converged one with the largest minimum weight.  Each
rule is then checked on its full point set against every spherical harmonic
up to its degree, and shown not to integrate degree L + 1 exactly.

The orbit structures are those of the published rules, and the solved rules
reproduce the published weights (tests/test_lebedev.py compares them).  For
the 350-point rule the search with the published structure (six b, two c,
three d orbits) did not converge, so it uses five b and three c orbits:
another rule of the same size and degree, with positive weights and other
points than the published one.

Run ``python lebedev_solve.py [sizes...]``; it rewrites lebedev_rules.inc
next to this file (minutes for the larger rules).  numpy and scipy only.
"""
from __future__ import annotations

import itertools
import sys
import time
from pathlib import Path

import numpy as np
from scipy.optimize import least_squares, nnls
from scipy.special import sph_harm_y

# npoints: (degree, orbit kinds, seed, weight projection, first try).  The
# larger rules converge from few starting points; "first try" skips (without
# solving) the starts of the seed's sequence that were found not to converge,
# so a run reproduces the table in minutes.
RULES = {
    50: (11, "a1 a2 a3 b", 1, "lstsq", 0),
    74: (13, "a1 a2 a3 b c", 1, "lstsq", 0),
    86: (15, "a1 a3 b b c", 1, "lstsq", 0),
    110: (17, "a1 a3 b b b c", 1, "lstsq", 0),
    170: (21, "a1 a2 a3 b b b c d", 1, "lstsq", 0),
    194: (23, "a1 a2 a3 b b b b c d", 1, "lstsq", 0),
    266: (27, "a1 a2 a3 b b b b b c d d", 1, "lstsq", 0),
    302: (29, "a1 a3 b b b b b b c c d d", 1, "lstsq", 0),
    350: (31, "a1 a3 b b b b b c c c d d d", 3, "nnls", 90),
    434: (35, "a1 a2 a3 b b b b b b b c c d d d d", 2, "lstsq", 17),
    590: (41, "a1 a3 b b b b b b b b b c c c d d d d d d", 5, "nnls", 12),
}
KIND = {"a1": 1, "a2": 2, "a3": 3, "b": 4, "c": 5, "d": 6}
NPTS = {"a1": 6, "a2": 12, "a3": 8, "b": 24, "c": 24, "d": 48}
NPRM = {"a1": 0, "a2": 0, "a3": 0, "b": 1, "c": 1, "d": 2}
PERMS = list(itertools.permutations(range(3)))
SIGNS = list(itertools.product((1.0, -1.0), repeat=3))


def images(p):
    """The 48 images of points p[R, 3] under the signed permutations."""
    return np.stack([p[:, pm] * np.array(s) for pm in PERMS for s in SIGNS], axis=1)


def re_ylm(l, m, pts):
    th = np.arccos(np.clip(pts[..., 2], -1, 1))
    ph = np.arctan2(pts[..., 1], pts[..., 0])
    return sph_harm_y(l, m, th, ph).real


def product_rule(L):
    """Gauss-Legendre x trapezoid rule, exact to degree 2L + 3 (weights sum to 1)."""
    t, wt = np.polynomial.legendre.leggauss(L + 2)
    nph = 2 * L + 4
    ph = 2 * np.pi * np.arange(nph) / nph
    st = np.sqrt(1 - t * t)
    p = np.stack([np.outer(st, np.cos(ph)), np.outer(st, np.sin(ph)), np.outer(t, np.ones(nph))], -1).reshape(-1, 3)
    w = (np.outer(wt, np.ones(nph)) * (2 * np.pi / nph)).reshape(-1) / (4 * np.pi)
    return p, w


class InvariantBasis:
    """Orthonormal O_h-invariant harmonics of even degree < L.  An invariant
    of degree l is a combination of Re Y_lm with m = 0 mod 4; the averages of
    those over O_h span the invariants, orthonormalised per degree."""

    def __init__(self, L):
        p, w = product_rule(L)
        img = images(p)
        self.terms = []
        for l in range(0, L, 2):
            ms = list(range(0, l + 1, 4))
            f = np.array([re_ylm(l, m, img).mean(axis=1) for m in ms])
            ev, vec = np.linalg.eigh((f * w) @ f.T)
            keep = ev > 1e-12
            if not keep.any():
                continue
            h = (vec[:, keep] / np.sqrt(ev[keep])).T @ f
            y = np.array([re_ylm(l, m, p) for m in ms])
            d = np.linalg.lstsq(y.T, h.T, rcond=None)[0].T
            assert np.abs(d @ y - h).max() < 1e-10
            self.terms.append((l, ms, d))
        self.count = sum(d.shape[0] for _, _, d in self.terms)

    def rows(self, lmax):
        return sum(d.shape[0] for l, _, d in self.terms if l < lmax)

    def eval(self, reps, lmax):
        return np.vstack([d @ np.array([re_ylm(l, m, reps) for m in ms]) for l, ms, d in self.terms if l < lmax])


def representatives(kinds, x):
    out, i = [], 0
    for k in kinds:
        if k == "a1":
            out.append((1.0, 0.0, 0.0))
        elif k == "a2":
            out.append((0.0, np.sqrt(0.5), np.sqrt(0.5)))
        elif k == "a3":
            out.append((np.sqrt(1 / 3),) * 3)
        elif k == "b":
            l = x[i]; i += 1
            out.append((l, l, np.sqrt(max(1 - 2 * l * l, 0.0))))
        elif k == "c":
            p = x[i]; i += 1
            out.append((p, np.sqrt(max(1 - p * p, 0.0)), 0.0))
        else:
            r, s = x[i], x[i + 1]; i += 2
            out.append((r, s, np.sqrt(max(1 - r * r - s * s, 0.0))))
    return np.array(out)


def parameter_bounds(kinds):
    lo, hi = [], []
    for k in kinds:
        if k == "b":
            lo.append(0.01); hi.append(0.705)
        elif k == "c":
            lo.append(0.708); hi.append(0.9999)
        elif k == "d":
            lo += [0.3, 0.02]; hi += [0.99, 0.705]
    return np.array(lo), np.array(hi)


def spread_start(kinds, rng):
    """Free coordinates spread along the edges (b, c) and over the interior (d)
    of the fundamental spherical triangle, with a random jitter."""
    nb, nc, nd = kinds.count("b"), kinds.count("c"), kinds.count("d")
    jitter = rng.uniform(0, 1)
    lb = np.linspace(0.0, np.sqrt(0.5), nb + 2)[1:-1] + jitter * rng.normal(size=nb) * 0.7 / (nb + 1)
    pc = np.cos(np.linspace(0, np.pi / 4, nc + 2)[1:-1]) + jitter * rng.normal(size=nc) * 0.3 / (nc + 1)
    xd = []
    for _ in range(nd):
        while True:
            v = np.sort(np.abs(rng.normal(size=3)))[::-1]
            v /= np.linalg.norm(v)
            if v[0] > v[1] > v[2] > 0.02:
                break
        xd += [v[0], v[1]]
    x, ib, ic, idd = [], 0, 0, 0
    for k in kinds:
        if k == "b":
            x.append(lb[ib]); ib += 1
        elif k == "c":
            x.append(pc[ic]); ic += 1
        elif k == "d":
            x += xd[2 * idd:2 * idd + 2]; idd += 1
    return np.array(x)


def weights_and_residual(kinds, basis, lmax, x, projection="lstsq"):
    """Weights by least squares (or nonnegative least squares, which steers
    the search of the larger rules towards positive weights) and the moment
    residual."""
    a = basis.eval(representatives(kinds, x), lmax) * np.array([NPTS[k] for k in kinds], float)
    b = np.zeros(a.shape[0])
    b[0] = 1.0  # the constant's moment; all others vanish
    w = nnls(a, b)[0] if projection == "nnls" else np.linalg.lstsq(a, b, rcond=None)[0]
    return w, a @ w - b


def solve(n, max_tries=2000, log=print):
    degree, spec, seed, projection, first = RULES[n]
    kinds = spec.split()
    basis = InvariantBasis(degree)
    lo, hi = parameter_bounds(kinds)
    assert basis.count == len(kinds) + len(lo), (basis.count, len(kinds) + len(lo))
    assert sum(NPTS[k] for k in kinds) == n
    stages = list(range(max(2, degree - 16), degree + 1, 4))
    if stages[-1] != degree + 1:
        stages.append(degree + 1)
    rng = np.random.default_rng(seed)
    best = None
    t0 = time.time()
    for t in range(max_tries):
        x = np.clip(spread_start(kinds, rng), lo + 1e-6, hi - 1e-6)
        if t < first:
            continue
        for lmax in stages:
            fun = lambda y, lmax=lmax: weights_and_residual(kinds, basis, lmax, y, projection)[1]
            x = least_squares(fun, x, bounds=(lo, hi), xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=200).x
        w, r = weights_and_residual(kinds, basis, degree + 1, x)
        if np.linalg.norm(r) < 1e-12:
            log(f"  {n}: try {t} converged, min weight {w.min():.6e} ({time.time() - t0:.0f} s)")
            if best is None or w.min() > best[1].min():
                best = (x, w)
            if w.min() > 0 or (n in (74, 266) and t >= 40):
                break
    if best is None:
        raise RuntimeError(f"no {n}-point rule found in {max_tries} tries")
    return kinds, degree, best[0], best[1]


def orbits_of(kinds, x, w):
    """(kind, a, b, weight) per orbit; orbits of a kind sorted by coordinates."""
    out, i = [], 0
    for k, wk in zip(kinds, w):
        a = b = 0.0
        if NPRM[k] == 1:
            a = x[i]; i += 1
        elif NPRM[k] == 2:
            a, b = x[i], x[i + 1]; i += 2
        out.append((KIND[k], float(a), float(b), float(wk)))
    return sorted(out, key=lambda o: (o[0], o[1], o[2]))


def points(orbits):
    pts, wts = [], []
    for kind, a, b, wk in orbits:
        if kind == 1:
            rep = (1.0, 0.0, 0.0)
        elif kind == 2:
            rep = (0.0, np.sqrt(0.5), np.sqrt(0.5))
        elif kind == 3:
            rep = (np.sqrt(1 / 3),) * 3
        elif kind == 4:
            rep = (a, a, np.sqrt(1 - 2 * a * a))
        elif kind == 5:
            rep = (a, np.sqrt(1 - a * a), 0.0)
        else:
            rep = (a, b, np.sqrt(1 - a * a - b * b))
        img = np.unique(np.round(images(np.array([rep]))[0], 15), axis=0)
        pts.append(img)
        wts.append(np.full(len(img), wk))
    return np.vstack(pts), np.concatenate(wts)


def check(n, degree, orbits):
    """Every harmonic up to the degree integrates exactly; degree + 1 does not."""
    p, w = points(orbits)
    assert len(p) == n, (n, len(p))
    worst = 0.0
    for l in range(degree + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), np.arccos(np.clip(p[:, 2], -1, 1)), np.arctan2(p[:, 1], p[:, 0]))
            y = y.real if m >= 0 else y.imag
            want = 1 / np.sqrt(4 * np.pi) if l == 0 else 0.0
            worst = max(worst, abs(w @ y - want))
    above = max(abs(w @ sph_harm_y(degree + 1, m, np.arccos(np.clip(p[:, 2], -1, 1)),
                                    np.arctan2(p[:, 1], p[:, 0])).real) for m in range(0, degree + 2))
    assert worst < 1e-13 and above > 1e-8, (n, worst, above)
    return worst, above


def write(rules, path):
    lines = ["// Lebedev-Laikov rules: orbit generators {kind, a, b, weight} (weights sum",
             "// to 1).  Written by lebedev_solve.py; do not edit.", ""]
    names = []
    for n in sorted(rules):
        degree, orbits = rules[n]
        name = f"kLebedev{n}"
        names.append((n, degree, len(orbits), name))
        lines.append(f"constexpr LebedevOrbit {name}[] = {{")
        for kind, a, b, wk in orbits:
            lines.append(f"    {{{kind}, {a!r}, {b!r}, {wk!r}}},")
        lines.append("};")
    lines.append("constexpr LebedevRule kLebedevRules[] = {")
    for n, degree, no, name in names:
        lines.append(f"    {{{n}, {degree}, {no}, {name}}},")
    lines.append("};")
    path.write_text("\n".join(lines) + "\n")


def main(argv):
    sizes = [int(a) for a in argv] or sorted(RULES)
    rules = {}
    for n in sizes:
        kinds, degree, x, w = solve(n)
        orbits = orbits_of(kinds, x, w)
        worst, above = check(n, degree, orbits)
        print(f"{n} points, degree {degree}: max moment error {worst:.1e} up to degree {degree}, "
              f"{above:.1e} at degree {degree + 1}; min weight {min(o[3] for o in orbits):.6e}", flush=True)
        rules[n] = (degree, orbits)
    write(rules, Path(__file__).resolve().parent / "lebedev_rules.inc")


if __name__ == "__main__":
    main(sys.argv[1:])
