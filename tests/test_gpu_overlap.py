"""Analytic AO overlap (the int1e_ovlp of the reference's overlap check,
python/isdf.py:665-671) against what fixes it independently:
* the closed-form overlap of contracted s Gaussians (H2 / STO-3G);
* unit diagonal of normalised spherical AOs;
* the grid quadrature of the oracle's AO values, which converges to the
  integral for every angular momentum, cartesian and spherical alike.
A wrong recursion term, transform row or AO index shows up at O(0.1); the
quadrature tolerance (measured level-4 grid error plus margin) is far below."""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


def test_s_overlap_closed_form(mp_isdf):
    mol = molecule.hydrogen_molecule("sto-3g")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    want = oracle.s_overlap(oracle.s_functions(mol.atm, mol.bas, mol.env))
    np.testing.assert_allclose(ctx.overlap(), want, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("basis,cart", [("cc-pvdz", False), ("6-31g*", True), ("6-31g*", False)])
def test_overlap_matches_grid_quadrature(mp_isdf, basis, cart):
    mol = molecule.single_water(basis, cart=cart)
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=cart)
    s = ctx.overlap()
    assert s.shape == (mol.nao, mol.nao)
    np.testing.assert_allclose(s, s.T, rtol=0, atol=1e-14)
    if not cart:
        np.testing.assert_allclose(np.diag(s), 1.0, rtol=0, atol=1e-10)
    xyz, w = gridmod.make_grid(mol, 4)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz, cart=cart)
    quad = ao.T @ (ao * w[:, None])
    err = np.abs(s - quad).max()
    print(f"{basis} cart={cart}: max |S - grid quadrature| = {err:.2e}")
    assert err < 1e-4
    # positive definite (a Gram matrix of independent functions)
    assert np.linalg.eigvalsh(s).min() > 0
