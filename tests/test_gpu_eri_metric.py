"""GPU integral-direct Cholesky Coulomb metric (mp_isdf.cholesky_eri,
mp_isdf.coulomb_metric) against the oracle's restatement of the reference
(cpp/src/cholesky_eri_greedy.cpp:53-175, cpp/include/cholesky_update.hpp:43-90,
cpp/include/cholesky_pivot.hpp:16-36, cpp/src/coulomb_metric.cpp:20-98), with
the reference tests' own checks (python/test/test_coulomb_metric.py:17-59,
:77-107, :109-133) restated on PySCF-free molecules: s shells (H2 6-31G*),
p and d shells (water cc-pVDZ, spherical; water 6-31G*, Cartesian d).

The GPU's ERIs sum the primitive pairs in another order than the oracle's
shell quartets, so the packed matrices agree to ~1e-15 relative and the
pivots agree wherever the residuals are not tied to rounding; the
distorted-water geometry keeps the hydrogens inequivalent."""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


def distorted_water(basis, cart=False):
    xyz = np.array([[0.0, 0.02, -0.01], [0.03, 0.77, 0.49], [-0.02, -0.73, 0.53]])
    return molecule.make_tables(["O", "H", "H"], xyz / molecule.BOHR, basis=basis, cart=cart)


def packed_full(packed, nao):
    mu, nu = np.tril_indices(nao)
    idx = np.zeros((nao, nao), dtype=int)
    idx[mu, nu] = np.arange(len(mu))
    idx[nu, mu] = np.arange(len(mu))
    return packed[idx[:, :, None, None], idx[None, None, :, :]]


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


@pytest.mark.parametrize("basis,cart", [("cc-pvdz", False), ("6-31g*", True), ("sto-3g", False)])
def test_cholesky_and_metric_match_oracle(mp_isdf, basis, cart):
    mol = distorted_water(basis, cart)
    nao = mol.nao
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=cart)
    packed = oracle.eri_packed(mol.atm, mol.bas, mol.env, cart=cart)
    npair = len(packed)
    # one step: the first pivot, its factor and the residual diagonal
    f1, p1, r1 = mp_isdf.cholesky_eri(ctx, tolerance=0, max_rank=1)
    assert f1.shape == (1, npair) and p1[0] == np.argmax(np.diag(packed))
    np.testing.assert_allclose(f1[0], packed[:, p1[0]] / np.sqrt(packed[p1[0], p1[0]]), atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(r1, np.maximum(0, np.diag(packed) - f1[0] ** 2), atol=1e-12, rtol=1e-12)
    # six greedy steps against the reference's own NumPy restatement
    expected, pivots, residual = reference_greedy(packed, 6)
    f6, p6, r6 = mp_isdf.cholesky_eri(ctx, tolerance=0, max_rank=6)
    np.testing.assert_array_equal(p6, pivots)
    np.testing.assert_allclose(f6, expected, atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(r6, residual, atol=1e-12, rtol=1e-12)
    # to tolerance 1e-12: the factorisation reproduces the packed matrix
    f, p, r = mp_isdf.cholesky_eri(ctx, tolerance=1e-12)
    fo, po, ro = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, tolerance=1e-12, cart=cart)
    m = min(len(p), len(po))
    differ = np.flatnonzero(p[:m] != po[:m])
    same = int(differ[0]) if len(differ) else m
    print(f"\nwater {basis}{' cart' if cart else ''}: npair {npair}, rank {len(p)} (oracle {len(po)}), "
          f"pivots equal for the first {same}")
    assert len(set(p.tolist())) == len(p)
    np.testing.assert_allclose(f.T @ f, packed, atol=2e-10, rtol=2e-10)
    np.testing.assert_allclose(r, np.maximum(0, np.diag(packed - f.T @ f)), atol=2e-10)
    assert r.max() <= 1e-12 * np.diag(packed).max() + 1e-12
    # run to run: bit-identical
    again = mp_isdf.cholesky_eri(ctx, tolerance=1e-12)
    for a, b in zip((f, p, r), again):
        np.testing.assert_array_equal(a, b)
    # the metric
    transform = np.random.default_rng(42).normal(size=(3, nao * nao))
    full = transform.reshape(3, nao, nao)
    want = np.einsum("aij,ijkl,bkl->ab", full, packed_full(packed, nao), full)
    for batch in (1, 2, 64):
        got = mp_isdf.coulomb_metric(ctx, transform, batch_size=batch)
        np.testing.assert_allclose(got, want, atol=2e-10, rtol=2e-10)
        np.testing.assert_allclose(got, got.T, atol=1e-12)
    mu, nu = np.tril_indices(nao)
    coeff = full[:, mu, nu].copy()
    coeff[:, mu != nu] += full[:, nu[mu != nu], mu[mu != nu]]
    proj = coeff @ f1[0]
    np.testing.assert_allclose(mp_isdf.coulomb_metric(ctx, transform, max_rank=1), np.outer(proj, proj), atol=1e-12)
    np.testing.assert_array_equal(mp_isdf.coulomb_metric(ctx, transform, tolerance=2), np.zeros((3, 3)))


def test_argument_errors(mp_isdf):
    mol = distorted_water("sto-3g")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    npair = mol.nao * (mol.nao + 1) // 2
    transform = np.zeros((2, mol.nao ** 2))
    with pytest.raises(ValueError):
        mp_isdf.coulomb_metric(ctx, transform, algorithm="dense")
    with pytest.raises(ValueError):
        mp_isdf.cholesky_eri(ctx, max_rank=npair + 1)
    with pytest.raises(ValueError):
        mp_isdf.coulomb_metric(ctx, transform, batch_size=0)
    with pytest.raises(ValueError):
        mp_isdf.cholesky_eri(ctx, tolerance=-1.0)
    with pytest.raises(ValueError):
        mp_isdf.coulomb_metric(ctx, np.zeros((2, mol.nao ** 2 + 1)))
    bad = transform.copy()
    bad[1, 3] = np.nan
    with pytest.raises(ValueError):
        mp_isdf.coulomb_metric(ctx, bad)
    np.testing.assert_array_equal(mp_isdf.coulomb_metric(ctx, np.zeros((0, mol.nao ** 2))), np.zeros((0, 0)))


def test_uniform_pivots_and_metric(mp_isdf):
    mol = molecule.hydrogen_molecule("6-31g*")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    packed = oracle.eri_packed(mol.atm, mol.bas, mol.env)
    kw = dict(algorithm="unblocked_uniform", seed=23)
    f, p, r = mp_isdf.cholesky_eri(ctx, **kw)
    for a, b in zip((f, p, r), mp_isdf.cholesky_eri(ctx, **kw)):
        np.testing.assert_array_equal(a, b)
    assert len(set(p.tolist())) == len(p)
    np.testing.assert_allclose(f.T @ f, packed, atol=2e-10, rtol=2e-10)
    assert r.max() <= 1e-12 * np.diag(packed).max() + 1e-12
    # the same draws as the oracle (std::mt19937_64 + uniform_int_distribution)
    _, po, _ = oracle.cholesky_eri(mol.atm, mol.bas, mol.env, **kw)
    np.testing.assert_array_equal(p, po)
    choices = {int(mp_isdf.cholesky_eri(ctx, max_rank=1, algorithm="unblocked_uniform", seed=s)[1][0])
               for s in range(16)}
    assert len(choices) > 1
    transform = np.random.default_rng(5).normal(size=(3, mol.nao ** 2))
    full = transform.reshape(3, mol.nao, mol.nao)
    want = np.einsum("aij,ijkl,bkl->ab", full, packed_full(packed, mol.nao), full)
    np.testing.assert_allclose(mp_isdf.coulomb_metric(ctx, transform, **kw), want, atol=2e-10, rtol=2e-10)


def test_water_dimer_metric_matches_oracle(mp_isdf):
    """(H2O)2 / cc-pVDZ (48 AOs, 1176 pairs): both factorisations to 1e-12
    give the same metric of three random transforms."""
    mol = molecule.water_cluster(2, "cc-pvdz")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    transform = np.random.default_rng(9).normal(size=(3, mol.nao ** 2))
    got = mp_isdf.coulomb_metric(ctx, transform)
    want = oracle.coulomb_metric(mol.atm, mol.bas, mol.env, transform)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"\n(H2O)2 cc-pVDZ metric: relative difference {err:.2e}")
    assert err <= 1e-10


def test_h2o8_cholesky_scale(mp_isdf):
    """(H2O)8 / cc-pVDZ (192 AOs, 18528 pairs) to 1e-8: the residual bound and
    the factor norms of the pivots, timed."""
    import time
    mol = molecule.water_cluster(8, "cc-pvdz")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    t = time.perf_counter()
    f, p, r = mp_isdf.cholesky_eri(ctx, tolerance=1e-8)
    dt = time.perf_counter() - t
    npair = f.shape[1]
    print(f"\n(H2O)8 cc-pVDZ: npair {npair}, rank {len(p)} at 1e-8 in {dt:.2f} s")
    assert len(set(p.tolist())) == len(p)
    d0 = np.max(np.sum(f * f, axis=0) + r)  # the initial diagonal's maximum
    assert r.max() <= 1e-8 * d0 * (1 + 1e-12)
    # the factor rows reproduce the pivot columns' diagonals exactly at the pivots
    diag = np.sum(f[:, p] ** 2, axis=0)
    assert np.all(diag > 0)


def test_driver_uses_ao_coefficients_and_anchor_weights(mp_isdf):
    """build_isdf's default metric (unblocked_greedy) on H2 / STO-3G, level 0,
    three anchors (python/test/test_coulomb_metric.py:156-170 restated): the
    metric is W_a T V T^T W_a with T = pinv(W P W P^T) W P, W = diag(w_anchor),
    P the anchor AO pairs, V the ERI tensor."""
    from mp_isdf_amd import isdf as driver
    mol = molecule.hydrogen_molecule("sto-3g")
    result = driver.build_isdf(mol, level=0, anchor_rank=3, compute_metric=True)
    anchor = result["X1"]
    weighted = (anchor[:, :, None] * anchor[:, None, :]).reshape(len(anchor), -1)
    weighted *= result["anchor_weights"][:, None]
    transform = np.linalg.pinv(weighted @ weighted.T, rcond=1e-12) @ weighted
    coeff = transform.reshape(len(anchor), mol.nao, mol.nao)
    eri = packed_full(oracle.eri_packed(mol.atm, mol.bas, mol.env), mol.nao)
    expected = np.einsum("aij,ijkl,bkl->ab", coeff, eri, coeff)
    w = result["anchor_weights"]
    np.testing.assert_allclose(result["coulomb_metric"], w[:, None] * expected * w[None, :], rtol=1e-10, atol=1e-10)
    assert result["timings"]["coulomb_metric"] > 0
