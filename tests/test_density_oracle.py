"""CPU: the oracle's density Khatri-Rao sketches (cpp/src/anchor_select_randomized_id.cpp:92-374,
policies AODensity and AODensityNormalized) against the reference's own
density-sketch test cases (python/test/test_density_sketch.py:35-121), which
pin them by closed-form outcomes: rank-one selections decided by mass and
AO norm, screening, validation, and equivalence with the quadrature sketch
under rescaled weights."""
import numpy as np

import oracle
import density_cases as cases


def oracle_select(mol):
    def select(points, weights, axes="ao", **kwargs):
        return oracle.select(mol.atm, mol.bas, mol.env, points, weights, axes, **kwargs)
    return select


def oracle_ao(mol):
    return lambda pts: oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)


def test_rank_one_cases():
    mol = cases.helium()
    cases.check_rank_one_cases(oracle_select(mol), oracle_ao(mol))


def test_screening_and_validation():
    mol = cases.helium()
    cases.check_screening_and_validation(oracle_select(mol), oracle_ao(mol))


def test_raw_matches_quadrature():
    cases.check_raw_matches_quadrature(oracle_select(cases.helium()))


def test_normalized_matches_rescaled_quadrature():
    mol = cases.hydrogen_pair()
    cases.check_normalized_matches_rescaled_quadrature(oracle_select(mol), oracle_ao(mol))


def test_density_sketch_identity():
    """test_density_sketch.py:123-144: the factored sketch of the normalized
    policy is the probe matrix applied to pair columns scaled by sqrt(m)/q."""
    rng = np.random.default_rng(41)
    phi = rng.normal(size=(7, 4))
    m = rng.uniform(.1, 2, 7)
    ell = 13
    u, v = rng.normal(size=(2, 4, ell))
    q = np.sum(phi ** 2, axis=1)
    sw = np.sqrt(m)
    got = oracle.sketch(phi / np.sqrt(q)[:, None], u, v, sw)
    pairs = np.einsum("gi,gj->gij", phi, phi).reshape(7, -1).T
    probes = np.einsum("it,jt->tij", u, v).reshape(ell, -1) / np.sqrt(ell)
    np.testing.assert_allclose(got.T, probes @ (pairs * (np.sqrt(m) / q)), atol=1e-14)
