"""Pins the oracle's selection algorithms (oracle/select_oracle.cpp) and the
NumPy driver maths (oracle/__init__.py) against the reference tests' own
oracles: scipy's dgeqp3 (scipy.linalg.qr(pivoting=True)), brute-force greedy
residuals, the Khatri-Rao identity and the driver's known-answer pair error.
"""
import numpy as np
import pytest
from scipy.linalg import qr

import oracle
from mp_isdf_amd import basis, molecule


@pytest.mark.parametrize("shape", [(30, 50), (50, 30), (7, 200), (1, 5)])
def test_geqp3_matches_scipy(shape):
    rng = np.random.default_rng(sum(shape))
    a = rng.normal(size=shape) * rng.uniform(0.1, 10, size=shape[1])
    f, piv, tau = oracle.geqp3(a)
    _, r, p2 = qr(a, pivoting=True, mode="economic")
    k = min(shape)
    np.testing.assert_array_equal(piv[:k], p2[:k])
    np.testing.assert_allclose(np.abs(np.diag(f)[:k]), np.abs(np.diag(r)), rtol=1e-12)


def test_geqp3_greedy_residual_brute_force():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(12, 40)) @ np.diag(rng.uniform(0.1, 3, 40))
    _, piv, _ = oracle.geqp3(a)
    for k in range(10):
        if k:
            q = np.linalg.qr(a[:, piv[:k]])[0]
            res = a - q @ (q.T @ a)
        else:
            res = a
        norms = np.linalg.norm(res, axis=0)
        assert np.isclose(norms[piv[k]], norms.max(), rtol=1e-12)


class H2Case:
    mol = molecule.hydrogen_molecule("sto-3g", half_bond_angstrom=0.5 * 0.529177210920)
    points = np.random.default_rng(12).normal(size=(8, 3))
    weights = np.array([1., 2., 0., -1., 1., 3., 1., 1.])


def reference_pivots(case, axes, tolerance=1e-10, max_rank=0, mo=None, nocc=0, axis_weights=None):
    """python/test/test_anchor_select.py:33-61 restated: explicit pair matrix + scipy."""
    good = case.weights != 0
    pts = case.points[good]
    m = case.mol
    ao = oracle.ao_eval(m.atm, m.bas, m.env, pts)
    grad = oracle.ao_eval(m.atm, m.bas, m.env, pts, deriv=1)
    pieces = {"ao": ao, "dx": grad[0], "dy": grad[1], "dz": grad[2],
              "occ": None if mo is None else ao @ mo[:, :nocc]}
    blocks = [pieces[a] for a in axes]
    if len(blocks) > 1:
        factors = axis_weights if axis_weights is not None else [1.] * len(blocks)
        blocks = [f * b / np.linalg.norm(b) for f, b in zip(factors, blocks)]
    cols = np.hstack(blocks)
    outer = (cols[:, :, None] * cols[:, None, :]).reshape(len(pts), -1)
    matrix = (outer * np.sqrt(np.abs(case.weights[good]))[:, None]).T
    _, r, piv = qr(matrix, pivoting=True, mode="economic")
    diag = np.abs(np.diag(r))
    count = int(np.sum(diag > tolerance * diag[0])) if len(diag) and diag[0] else 0
    if max_rank:
        count = min(count, max_rank)
    return np.flatnonzero(good)[piv[:count]].tolist()


@pytest.mark.parametrize("axes,max_rank", [(["ao"], 0), (["dx", "dz"], 4), (["dz", "ao", "dz"], 3),
                                           (["ao", "dx"], 3)])
def test_dense_select_matches_scipy(axes, max_rank):
    c = H2Case
    got = oracle.select(c.mol.atm, c.mol.bas, c.mol.env, c.points, c.weights, axes, max_rank=max_rank)
    assert got == reference_pivots(c, axes, max_rank=max_rank)
    assert 2 not in got


def test_dense_select_occupied_and_axis_weights():
    c = H2Case
    mo = np.array([[0.55, 1.2], [0.55, -1.2]])
    got = oracle.select(c.mol.atm, c.mol.bas, c.mol.env, c.points, c.weights, ["ao", "occ"], max_rank=3,
                        mo_coeff=mo, nocc=1, axis_weights=[1., 2.])
    assert got == reference_pivots(c, ["ao", "occ"], max_rank=3, mo=mo, nocc=1, axis_weights=[1., 2.])


class ThreeS:
    atm, bas, env = basis.three_s_shell_tables()
    points = np.random.default_rng(23).normal(size=(17, 3))
    weights = np.linspace(0.5, 2, 17)
    weights[2] = 0
    weights[4] *= -1


@pytest.mark.parametrize("batch", [1, 3, 17])
def test_online_matches_scipy_and_greedy(batch):
    """python/test/test_online_anchor_select.py:44-69."""
    c = ThreeS
    mask = c.weights != 0
    f = oracle.axis_eval(c.atm, c.bas, c.env, c.points[mask], ["ao"])
    matrix = (np.einsum("gi,gj->gij", f, f).reshape(mask.sum(), -1)
              * np.sqrt(np.abs(c.weights[mask]))[:, None]).T
    expected = np.flatnonzero(mask)[qr(matrix, pivoting=True)[2]]
    got = oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", algorithm="qrcp_online",
                        batch_size=batch)
    assert len(got) == 6
    np.testing.assert_array_equal(got, expected[:6])
    got3 = oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", algorithm="qrcp_online",
                         batch_size=batch, max_rank=3)
    np.testing.assert_array_equal(got3, expected[:3])


def test_online_cutoffs_equal_dense():
    c = ThreeS
    assert oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", algorithm="qrcp_online",
                         tolerance=1) == []
    for cutoff in (0.1, 0.5):
        assert (oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", algorithm="qrcp_online",
                              tolerance=cutoff, batch_size=2)
                == oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", tolerance=cutoff))


def test_randomized_rules():
    """python/test/test_randomized_anchor_select.py:46-67 (counts and rank rules)."""
    c = ThreeS
    w = np.linspace(.4, 1.5, 19)
    w[2] = 0
    w[4] = -1
    pts = np.random.default_rng(23).normal(size=(19, 3))
    kw = dict(algorithm="rand_sketch", sketch_size=12, oversampling=3, seed=31)
    piv = oracle.select(c.atm, c.bas, c.env, pts, w, "ao", max_rank=3, batch_size=4, **kw)
    assert len(piv) == 3 and 2 not in piv
    assert piv == oracle.select(c.atm, c.bas, c.env, pts, w, "ao", max_rank=3, batch_size=2, **kw)
    assert len(oracle.select(c.atm, c.bas, c.env, pts, w, "ao", batch_size=4, **kw)) == 6
    assert oracle.select(c.atm, c.bas, c.env, pts, w, "ao", tolerance=1, **kw) == []
    assert len(oracle.select(c.atm, c.bas, c.env, pts, w, "ao", tolerance=1, max_rank=2, **kw)) == 2
    with pytest.raises(ValueError):
        oracle.select(c.atm, c.bas, c.env, pts, w, "ao", max_rank=7, **kw)


def test_khatri_rao_identity():
    """python/test/test_randomized_anchor_select.py:110-119 with the oracle's sketch."""
    rng = np.random.default_rng(7)
    phi = rng.normal(size=(9, 4))
    w = rng.uniform(.1, 2., size=9)
    u, v = oracle.omega(4, 6, seed=3)
    y = oracle.sketch(phi, u, v, np.sqrt(w))
    pairs = np.einsum("gi,gj->gij", phi, phi).reshape(9, 16)
    projection = np.einsum("is,js->sij", u, v).reshape(6, 16) / np.sqrt(6)
    np.testing.assert_allclose(y, (projection @ (pairs * np.sqrt(w)[:, None]).T).T, rtol=1e-12)


def test_omega_is_standard_normal_and_seeded():
    u, v = oracle.omega(64, 500, seed=7)
    u2, _ = oracle.omega(64, 500, seed=7)
    np.testing.assert_array_equal(u, u2)
    for x in (u, v):
        assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    assert abs(np.corrcoef(u.ravel(), v.ravel())[0, 1]) < 0.01


def test_pair_error_known_answer():
    """python/test/test_isdf.py:157-164."""
    err = oracle.pair_space_error_norms(np.array([[1., 2.], [3., 4.]]), np.array([[1., 0.]]),
                                        np.array([1.]), np.array([[1., 3.]]), np.array([2., -1.]),
                                        block_size=1)
    np.testing.assert_allclose(err["l2"], np.sqrt(24.))
    np.testing.assert_allclose(err["relative"], np.sqrt(24. / 35.))


def test_refit_matches_lstsq():
    """python/test/test_isdf.py:41-49 with the oracle's refit."""
    mol = molecule.hydrogen_molecule("sto-3g")
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(60, 3))
    wts = rng.uniform(0.2, 1.0, 60)
    anchors = oracle.select(mol.atm, mol.bas, mol.env, pts, wts, "ao", max_rank=3)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    aa = ao[anchors]
    pairs = (aa[:, :, None] * aa[:, None, :]).reshape(len(anchors), -1)
    theta, _, diag = oracle.physical_pair_refit(pairs, wts[anchors], ao, 7)
    products = (ao[:, :, None] * ao[:, None, :]).reshape(len(ao), -1)
    physical = np.linalg.lstsq(pairs.T, products.T, rcond=1e-12)[0]
    np.testing.assert_allclose(theta, physical / wts[anchors][:, None], rtol=1e-12, atol=1e-12)
    stream = oracle.stream_pair_space_error_norms(ao, ao, pairs, wts[anchors], theta, wts)
    block = oracle.pair_space_error_norms(products, pairs, wts[anchors], theta, wts)
    np.testing.assert_allclose(stream["l2"], block["l2"], rtol=1e-12, atol=1e-14)
    assert stream["relative"] < 1e-10  # 3 anchors span the 3 symmetric STO-3G H2 pairs
