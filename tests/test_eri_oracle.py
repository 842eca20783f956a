"""Pins of the oracle's integral-direct Cholesky Coulomb metric
(oracle/eri_oracle.cpp), without PySCF or libcint:

* the Boys function against its defining integral (scipy quadrature) and
  F_n(0) = 1 / (2n + 1);
* s-shell ERIs against the closed form (Boys 1950) of oracle.s_eri, itself
  pinned by Szabo and Ostlund's H2/STO-3G integrals;
* p and d primitives against the translational-derivative relations
  (x - A_x) g = (1 / 2a) dg/dA_x,  (x - A_x)^2 g = (1 / 4a^2) d^2g/dA_x^2 + g / 2a,
  (x - A_x)(y - A_y) g = (1 / 4a^2) d^2g/dA_x dA_y, on every one of the four
  centres, by high-order central differences of the s-type ERI;
* the contracted, transformed AO ERIs of water (s, p, spherical and Cartesian
  d) against the grid-quadrature Coulomb metric of the AO pair densities;
* the 8-fold symmetry and positive semidefiniteness of the packed matrix;
* the greedy and uniform Cholesky against the reference tests' own NumPy
  restatement (python/test/test_coulomb_metric.py:62-78, :17-59, :77-107).
"""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import molecule


def test_boys_function():
    from scipy.integrate import quad
    np.testing.assert_allclose(oracle.boys(8, 0.0), 1.0 / (2 * np.arange(9) + 1), rtol=1e-15)
    for t in (1e-9, 0.3, 2.5, 11.0, 29.999, 30.001, 47.0, 120.0):
        f = oracle.boys(8, t)
        for n in range(9):
            ref = quad(lambda u: u ** (2 * n) * np.exp(-t * u * u), 0, 1, epsabs=0, epsrel=1e-13, limit=200)[0]
            np.testing.assert_allclose(f[n], ref, rtol=5e-13, err_msg=f"F_{n}({t})")


def _packed_full(packed, nao):
    """(nao, nao, nao, nao) tensor from the packed matrix."""
    mu, nu = np.tril_indices(nao)
    idx = np.zeros((nao, nao), dtype=int)
    idx[mu, nu] = np.arange(len(mu))
    idx[nu, mu] = np.arange(len(mu))
    return packed[idx[:, :, None, None], idx[None, None, :, :]]


def test_s_shells_match_closed_form():
    for mol in (molecule.hydrogen_molecule("sto-3g"), molecule.hydrogen_molecule("6-31g*")):
        got = oracle.eri_packed(mol.atm, mol.bas, mol.env)
        full = oracle.s_eri(oracle.s_functions(mol.atm, mol.bas, mol.env))
        np.testing.assert_allclose(_packed_full(got, mol.nao), full, rtol=1e-13, atol=1e-15)


def _shift(centres, slot, axis, h):
    c = centres.copy()
    c[slot, axis] += h
    return c


def _d1(f, h):
    return (f(-2 * h) - 8 * f(-h) + 8 * f(h) - f(2 * h)) / (12 * h)


def _d2(f, h):
    return (-f(2 * h) + 16 * f(h) - 30 * f(0.0) + 16 * f(-h) - f(-2 * h)) / (12 * h * h)


@pytest.mark.parametrize("slot", [0, 1, 2, 3])
def test_p_and_d_primitives_from_derivatives(slot):
    rng = np.random.default_rng(10 + slot)
    exps = rng.uniform(0.4, 1.6, size=4)
    centres = rng.normal(scale=0.8, size=(4, 3))
    s = np.zeros((4, 3), dtype=int)
    a = exps[slot]

    def s_eri(dx=0.0, dy=0.0, axis=0, axis2=1):
        return oracle.eri_prim(s, exps, _shift(_shift(centres, slot, axis, dx), slot, axis2, dy))

    for axis in range(3):
        pw = s.copy()
        pw[slot, axis] = 1
        want = _d1(lambda h: s_eri(h, 0.0, axis), 1e-3) / (2 * a)
        np.testing.assert_allclose(oracle.eri_prim(pw, exps, centres), want, rtol=2e-9, atol=1e-12)
        pw[slot, axis] = 2
        want = _d2(lambda h: s_eri(h, 0.0, axis), 2e-3) / (4 * a * a) + s_eri() / (2 * a)
        np.testing.assert_allclose(oracle.eri_prim(pw, exps, centres), want, rtol=2e-7, atol=1e-10)
    pw = s.copy()
    pw[slot, 0] = pw[slot, 1] = 1
    h = 1e-3
    want = (s_eri(h, h) - s_eri(h, -h) - s_eri(-h, h) + s_eri(-h, -h)) / (4 * h * h) / (4 * a * a)
    np.testing.assert_allclose(oracle.eri_prim(pw, exps, centres), want, rtol=1e-5, atol=1e-9)


def test_mixed_centres_from_derivatives():
    """(p_x s | p_y p_z): one derivative on each of three centres."""
    rng = np.random.default_rng(4)
    exps = rng.uniform(0.5, 1.5, size=4)
    centres = rng.normal(scale=0.7, size=(4, 3))
    s = np.zeros((4, 3), dtype=int)

    def f(moves):
        c = centres.copy()
        for (slot, axis), h in moves.items():
            c[slot, axis] += h
        return oracle.eri_prim(s, exps, c)

    h = 2e-3
    keys = [(0, 0), (2, 1), (3, 2)]
    total = 0.0
    for signs in np.ndindex(2, 2, 2):
        sg = [1 if b == 0 else -1 for b in signs]
        total += np.prod(sg) * f({k: g * h for k, g in zip(keys, sg)})
    want = total / (2 * h) ** 3 / (2 * exps[0]) / (2 * exps[2]) / (2 * exps[3])
    pw = s.copy()
    pw[0, 0] = pw[2, 1] = pw[3, 2] = 1
    np.testing.assert_allclose(oracle.eri_prim(pw, exps, centres), want, rtol=2e-5)


@pytest.mark.parametrize("basis,cart", [("sto-3g", False), ("cc-pvdz", False), ("6-31g*", True)])
def test_packed_matrix_symmetry_and_grid_quadrature(basis, cart):
    from mp_isdf_amd import grid as gridmod
    mol = molecule.single_water(basis, cart=cart)
    packed = oracle.eri_packed(mol.atm, mol.bas, mol.env, cart=cart)
    np.testing.assert_allclose(packed, packed.T, rtol=1e-12, atol=1e-14)
    ev = np.linalg.eigvalsh(packed)
    assert ev.min() >= -1e-12 * ev.max()
    # the grid-quadrature Coulomb metric of the AO pair densities (pinned by
    # the Gaussian self-energy) on a level-3 grid resolves the valence pairs
    xyz, w = gridmod.make_grid(mol, 3)
    keep = w != 0
    xyz, w = xyz[keep], w[keep]
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz, cart=cart)
    mu, nu = np.tril_indices(mol.nao)
    # pairs of functions whose tightest primitive exponent is below 4
    rng = np.random.default_rng(1)
    exps = _ao_max_exponents(mol, cart)
    soft = np.flatnonzero((exps[mu] < 4.0) & (exps[nu] < 4.0))
    pick = rng.choice(soft, size=min(12, len(soft)), replace=False)
    dens = (ao[:, mu[pick]] * ao[:, nu[pick]]).T
    quad = oracle.grid_coulomb_metric(dens, xyz, w)
    ref = packed[np.ix_(pick, pick)]
    err = np.abs(quad - ref).max() / np.abs(ref).max()
    print(f"\n{basis}: {len(pick)} valence pairs, grid quadrature vs analytic ERIs {err:.2e}")
    assert err < 1e-2


def _ao_max_exponents(mol, cart):
    """Largest primitive exponent with a nonzero coefficient in each AO's contraction."""
    out = []
    for row in np.asarray(mol.bas):
        _, l, nprim, nctr, _, pexp, pcoef = (int(v) for v in row[:7])
        exps = mol.env[pexp:pexp + nprim]
        ncomp = (l + 1) * (l + 2) // 2 if cart else 2 * l + 1
        for c in range(nctr):
            coef = mol.env[pcoef + c * nprim:pcoef + (c + 1) * nprim]
            out += [exps[coef != 0].max()] * ncomp
    return np.asarray(out)


def reference_greedy(matrix, max_rank):
    """python/test/test_coulomb_metric.py:62-78, restated."""
    residual = np.maximum(0.0, np.diag(matrix).copy())
    factors, pivots = [], []
    for _ in range(max_rank):
        pivot = int(np.argmax(residual))
        maximum = residual[pivot]
        if maximum <= 0:
            break
        column = matrix[:, pivot].copy()
        for factor in factors:
            column -= factor * factor[pivot]
        column /= np.sqrt(maximum)
        factors.append(column)
        pivots.append(pivot)
        residual = np.maximum(0.0, residual - column ** 2)
        residual[pivot] = 0.0
    return np.array(factors), pivots, residual


def distorted_water(basis, cart=False):
    xyz = np.array([[0.0, 0.02, -0.01], [0.03, 0.77, 0.49], [-0.02, -0.73, 0.53]])
    return molecule.make_tables(["O", "H", "H"], xyz / molecule.BOHR, basis=basis, cart=cart)


@pytest.mark.parametrize("basis", ["sto-3g", "cc-pvdz"])
def test_greedy_cholesky_and_metric(basis):
    mol = distorted_water(basis)
    nao = mol.nao
    packed = oracle.eri_packed(mol.atm, mol.bas, mol.env)
    npair = len(packed)
    f1, p1, r1 = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, tolerance=0, max_rank=1)
    assert f1.shape == (1, npair) and p1[0] == np.argmax(np.diag(packed))
    np.testing.assert_allclose(f1[0], packed[:, p1[0]] / np.sqrt(packed[p1[0], p1[0]]), atol=1e-12)
    np.testing.assert_allclose(r1, np.maximum(0, np.diag(packed) - f1[0] ** 2), atol=1e-12)
    expected, pivots, residual = reference_greedy(packed, 6)
    f6, p6, r6 = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, tolerance=0, max_rank=6)
    np.testing.assert_array_equal(p6, pivots)
    np.testing.assert_allclose(f6, expected, atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(r6, residual, atol=1e-12, rtol=1e-12)
    f, p, r = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, tolerance=1e-12)
    assert len(set(p.tolist())) == len(p)
    np.testing.assert_allclose(f.T @ f, packed, atol=2e-10, rtol=2e-10)
    np.testing.assert_allclose(r, np.maximum(0, np.diag(packed - f.T @ f)), atol=2e-10)
    assert r.max() <= 1e-12 * np.diag(packed).max() + 1e-12
    transform = np.random.default_rng(42).normal(size=(3, nao * nao))
    full = transform.reshape(3, nao, nao)
    want = np.einsum("aij,ijkl,bkl->ab", full, _packed_full(packed, nao), full)
    for batch in (1, 2, 64):
        np.testing.assert_allclose(oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, batch_size=batch),
                                   want, atol=2e-10, rtol=2e-10)
    mu, nu = np.tril_indices(nao)
    coeff = full[:, mu, nu].copy()
    coeff[:, mu != nu] += full[:, nu[mu != nu], mu[mu != nu]]
    proj = coeff @ f1[0]
    np.testing.assert_allclose(oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, max_rank=1),
                               np.outer(proj, proj), atol=1e-12)
    np.testing.assert_array_equal(oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, tolerance=2),
                                  np.zeros((3, 3)))
    with pytest.raises(ValueError):
        oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, algorithm="dense")
    with pytest.raises(ValueError):
        oracle.cholesky_eri(mol.atm, mol.bas, mol.env, max_rank=npair + 1)
    with pytest.raises(ValueError):
        oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, batch_size=0)
    with pytest.raises(ValueError):
        oracle.cholesky_eri(mol.atm, mol.bas, mol.env, tolerance=-1.0)


def test_uniform_pivots_and_metric():
    mol = molecule.hydrogen_molecule("6-31g*")
    packed = oracle.eri_packed(mol.atm, mol.bas, mol.env)
    kw = dict(algorithm="unblocked_uniform", seed=23)
    first = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, **kw)
    second = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, **kw)
    for a, b in zip(first, second):
        np.testing.assert_array_equal(a, b)
    f, p, r = first
    assert len(set(p.tolist())) == len(p)
    np.testing.assert_allclose(f.T @ f, packed, atol=2e-10, rtol=2e-10)
    assert r.max() <= 1e-12 * np.diag(packed).max() + 1e-12
    choices = {int(oracle.cholesky_eri(mol.atm, mol.bas, mol.env, max_rank=1, algorithm="unblocked_uniform",
                                       seed=s)[1][0]) for s in range(16)}
    assert len(choices) > 1
    transform = np.random.default_rng(5).normal(size=(3, mol.nao ** 2))
    full = transform.reshape(3, mol.nao, mol.nao)
    want = np.einsum("aij,ijkl,bkl->ab", full, _packed_full(packed, mol.nao), full)
    np.testing.assert_allclose(oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform, **kw), want,
                               atol=2e-10, rtol=2e-10)
