"""Lebedev-Laikov angular rules (the angular part of PySCF's dft.gen_grid,
which the reference takes its grid from: python/isdf.py:488-496).

Pinned by what the mathematics fixes, not by the solver that wrote the table:
every spherical harmonic up to the rule's degree integrates exactly and one
of degree + 1 does not; the point set is closed under the 48 signed
permutations of O_h; and the 50-point rule equals Lebedev's closed form
(weights 4/315, 64/2835, 27/1280, 14641/725760, b orbit at l = 1/sqrt(11))."""
import itertools
import importlib.util
from pathlib import Path

import numpy as np
import pytest
from scipy.special import sph_harm_y

import mp_isdf_amd

DEGREE = {50: 11, 74: 13, 86: 15, 110: 17, 170: 21, 194: 23, 266: 27, 302: 29, 350: 31, 434: 35, 590: 41}


def rule(n):
    return mp_isdf_amd.mp_isdf.lebedev(n)


def tabulated():
    out = []
    for n in DEGREE:
        try:
            rule(n)
            out.append(n)
        except ValueError:
            pass
    return out


def moments(x, w, l):
    th = np.arccos(np.clip(x[:, 2], -1, 1))
    ph = np.arctan2(x[:, 1], x[:, 0])
    return np.array([w @ sph_harm_y(l, m, th, ph) for m in range(0, l + 1)])


def test_levels_zero_to_three_are_tabulated():
    for n in (50, 74, 86, 110, 170, 194, 266, 302):
        assert n in tabulated()


@pytest.mark.parametrize("n", sorted(DEGREE))
def test_rule_is_exact_to_its_degree(n):
    if n not in tabulated():
        pytest.skip(f"{n}-point rule not tabulated")
    x, w = rule(n)
    assert x.shape == (n, 3) and w.shape == (n,)
    np.testing.assert_allclose(np.linalg.norm(x, axis=1), 1.0, atol=1e-15)
    assert len(np.unique(np.round(x, 12), axis=0)) == n
    L = DEGREE[n]
    for l in range(L + 1):
        want = np.zeros(l + 1)
        if l == 0:
            want[0] = 1 / np.sqrt(4 * np.pi)
        assert np.abs(moments(x, w, l) - want).max() < 2e-15, (n, l)
    assert np.abs(moments(x, w, L + 1)).max() > 1e-6  # not exact one degree higher
    # closed under O_h: every signed permutation maps the rule onto itself
    key = {tuple(np.round(p, 12)): wi for p, wi in zip(x, w)}
    for perm in itertools.permutations(range(3)):
        for s in itertools.product((1, -1), repeat=3):
            for p, wi in zip(x[:, perm] * np.array(s), w):
                assert abs(key[tuple(np.round(p + 0.0, 12))] - wi) < 1e-16


def test_fifty_point_rule_closed_form():
    x, w = rule(50)
    r = np.sort(np.abs(x), axis=1)
    kinds = {}
    for p, wi in zip(r, w):
        kinds.setdefault(tuple(np.round(p, 12)), set()).add(round(wi, 15))
    l = 1 / np.sqrt(11)
    want = {(0.0, 0.0, 1.0): 4 / 315, (0.0, np.sqrt(0.5), np.sqrt(0.5)): 64 / 2835,
            (np.sqrt(1 / 3),) * 3: 27 / 1280, (l, l, np.sqrt(1 - 2 * l * l)): 14641 / 725760}
    assert len(kinds) == 4
    for rep, wi in want.items():
        got = kinds[tuple(np.round(np.sort(rep), 12))]
        assert len(got) == 1 and abs(got.pop() - wi) < 1e-15


# Orbit weights of Lebedev and Laikov's published tables (a1, a2, a3 orbits),
# which the solved rules must reproduce: exactness alone does not single out
# the published rule, these do.
PUBLISHED = {
    74: {1: 5.130671797338464e-4, 2: 1.660406956574204e-2, 3: -2.958603896103896e-2},
    86: {1: 1.154401154401154e-2, 3: 1.194390908585628e-2},
    110: {1: 3.828270494937162e-3},
    170: {1: 5.544842902037365e-3},
    194: {1: 1.782340447244611e-3},
    266: {1: -1.313769127326952e-3, 2: -2.522728704859336e-3, 3: 4.186853881700583e-3},
    302: {1: 8.545911725128148e-4, 3: 3.599119285025571e-3},
    434: {1: 5.265897968224436e-4},
    590: {1: 3.095121295306187e-4},
}
VERTEX = {1: (0.0, 0.0, 1.0), 2: (0.0, np.sqrt(0.5), np.sqrt(0.5)), 3: (np.sqrt(1 / 3),) * 3}


@pytest.mark.parametrize("n", sorted(PUBLISHED))
def test_rule_is_the_published_one(n):
    if n not in tabulated():
        pytest.skip(f"{n}-point rule not tabulated")
    x, w = rule(n)
    r = np.sort(np.abs(x), axis=1)
    for kind, want in PUBLISHED[n].items():
        at = np.all(np.abs(r - np.array(VERTEX[kind])) < 1e-12, axis=1)
        assert at.sum() == {1: 6, 2: 12, 3: 8}[kind]
        np.testing.assert_allclose(w[at], want, rtol=1e-11)


def test_unknown_size_raises():
    with pytest.raises(ValueError, match="Lebedev"):
        rule(51)


def test_solver_reproduces_the_small_rules():
    """lebedev_solve.py, which wrote the table, still finds the same rules."""
    path = Path(mp_isdf_amd.__file__).parent / "csrc" / "lebedev_solve.py"
    spec = importlib.util.spec_from_file_location("lebedev_solve", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    for n in (50, 86, 110):
        kinds, degree, xs, ws = mod.solve(n, log=lambda *a: None)
        orbits = mod.orbits_of(kinds, xs, ws)
        mod.check(n, degree, orbits)
        x, w = rule(n)
        mine = sorted(np.round(w, 13))
        theirs = sorted(np.round(np.concatenate([np.full(mod.NPTS[k], o[3]) for k, o in
                                                 zip(sorted(kinds, key=lambda k: mod.KIND[k]), orbits)]), 13))
        assert mine == theirs


def test_grid_uses_pyscf_sizes():
    """Level 3: 302-point shells pruned to [50, 86, 266, 302, 266] by
    r / R_Bragg; per water 75 O and 2 x 50 H radial shells."""
    from mp_isdf_amd import grid as gridmod, molecule
    mol = molecule.single_water("sto-3g")
    xyz, w = gridmod.make_grid(mol, 3, device=False)
    npts = len(w)
    assert npts > 75 * 50 and npts < 3 * 75 * 302
    # the integral of a normalised Gaussian density on each atom
    for a in range(3):
        r2 = np.sum((xyz - mol.coords[a]) ** 2, axis=1)
        alpha = 1.3
        val = w @ ((alpha / np.pi) ** 1.5 * np.exp(-alpha * r2))
        assert abs(val - 1) < 1e-6, (a, val)
