"""Coulomb energy of the fitted density and its nuclear adjoint (GPU).

E_J = tr(J D) / 2 for the ISDF J of evaluate_isdf_jk (python/isdf.py:130-151)
with the grid Coulomb metric of the refit (isdf.py:48-67, :238-263, :674-689).
The GPU evaluates it as u^T W K W u / 2 (u the fitted density on the grid)
without theta or the metric:

* the energy equals the oracle's chain (SVD refit, NumPy grid metric,
  evaluate_isdf_jk, tr(J D) / 2) and the product's own theta/metric/J chain;
* dE_J/dR from the hand-written adjoint equals central differences in FP64
  (h = 1e-4 Bohr, grid and anchor points fixed) to 1e-6 of max |dE_J/dR|;
* the torch.autograd.Function returns the same gradient through backward().
"""
import numpy as np
import pytest
import torch

import oracle
from mp_isdf_amd import adjoint, molecule

pytestmark = pytest.mark.gpu

H = 1e-4
FD_RTOL = 1e-6


def density_matrix(nao, nocc=5, seed=3):
    """A fixed symmetric positive semidefinite AO density matrix 2 B B^T."""
    b = np.random.default_rng(seed).normal(scale=0.3, size=(nao, nocc))
    return 2.0 * b @ b.T


def small_case(basis="6-31g*", cart=True, npts=1500, nanchor=40, seed=5):
    mol = molecule.single_water(basis, cart=cart)
    rng = np.random.default_rng(seed)
    pts = mol.coords[rng.integers(0, mol.natm, npts)] + rng.normal(scale=1.1, size=(npts, 3))
    w = rng.uniform(0.05, 1.0, size=npts)
    anchors = np.asarray(oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", max_rank=nanchor, cart=cart))
    return mol, pts, w, anchors, density_matrix(mol.nao)


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_energy_matches_oracle_chain(mp_isdf, basis, cart):
    mol, pts, w, anchors, dm = small_case(basis, cart)
    e = adjoint.coulomb_energy_value(mol, pts[anchors], pts, w, dm, batch_size=256)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, cart=cart)
    xa = ao[anchors]
    pairs = (xa[:, :, None] * xa[:, None, :]).reshape(len(anchors), -1)
    theta, _, _ = oracle.physical_pair_refit(pairs, w[anchors], ao, 256)
    z = w[anchors][:, None] * oracle.grid_coulomb_metric(theta, pts, w) * w[anchors][None, :]
    j, _ = oracle.evaluate_isdf_jk(xa, z, dm)
    want = 0.5 * np.sum(j * dm)
    print(f"\n{basis}: E_J {e:.12e} oracle {want:.12e}")
    np.testing.assert_allclose(e, want, rtol=1e-9)
    # the product's own chain: theta from the GPU fit, GPU metric, GPU J
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=cart)
    fit = mp_isdf.isdf_fit(ctx, pts[anchors], w[anchors], pts, w, interp_vecs=True)
    zg = w[anchors][:, None] * mp_isdf.grid_coulomb_metric(fit["interp_vecs"], pts, w) * w[anchors][None, :]
    jg, _ = mp_isdf.isdf_jk(np.ascontiguousarray(xa), zg, dm)
    np.testing.assert_allclose(e, 0.5 * np.sum(jg * dm), rtol=1e-9)


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_energy_gradient_matches_central_differences(mp_isdf, basis, cart):
    mol, pts, w, anchors, dm = small_case(basis, cart)
    e, grad, gao = adjoint.coulomb_energy_and_gradient(mol, pts[anchors], pts, w, dm, batch_size=256)
    assert grad.shape == (mol.natm, 3) and gao.shape == (3, mol.nao)
    np.testing.assert_allclose(e, adjoint.coulomb_energy_value(mol, pts[anchors], pts, w, dm), rtol=1e-11)
    fd = np.zeros_like(grad)
    for a in range(mol.natm):
        for c in range(3):
            vals = []
            for sgn in (1, -1):
                x = mol.coords.copy()
                x[a, c] += sgn * H
                vals.append(adjoint.coulomb_energy_value(adjoint.with_coords(mol, x), pts[anchors], pts, w, dm))
            fd[a, c] = (vals[0] - vals[1]) / (2 * H)
    scale = np.abs(grad).max()
    print(f"\n{basis}: max |dE_J/dR| {scale:.6e}, max |adjoint - FD| {np.abs(grad - fd).max() / scale:.3e} relative")
    assert scale > 0
    assert np.abs(grad - fd).max() <= FD_RTOL * scale, (grad, fd)
    # translation invariance would hold only if the grid moved with the atoms;
    # with a fixed grid the forces need not sum to zero, so nothing is asserted.


def test_energy_autograd_function_backward(mp_isdf):
    mol, pts, w, anchors, dm = small_case()
    coords = torch.tensor(mol.coords, dtype=torch.float64, requires_grad=True)
    e = adjoint.coulomb_energy(coords, mol, pts[anchors], pts, w, dm, batch_size=256)
    e.backward()
    _, grad, _ = adjoint.coulomb_energy_and_gradient(mol, pts[anchors], pts, w, dm, batch_size=256)
    np.testing.assert_allclose(coords.grad.numpy(), grad, rtol=1e-10, atol=1e-14)


def test_energy_argument_errors(mp_isdf):
    mol, pts, w, anchors, dm = small_case("sto-3g", False, npts=300, nanchor=10)
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    with pytest.raises(ValueError, match="symmetric"):
        bad = dm.copy()
        bad[0, 1] += 1.0
        mp_isdf.coulomb_energy_gradient(ctx, pts[anchors], pts, w, bad)
    with pytest.raises(ValueError, match="nonzero"):
        w0 = w.copy()
        w0[3] = 0.0
        mp_isdf.coulomb_energy_gradient(ctx, pts[anchors], pts, w0, dm)
    with pytest.raises(ValueError, match="shape"):
        mp_isdf.coulomb_energy_gradient(ctx, pts[anchors], pts, w, dm[:-1, :-1])
