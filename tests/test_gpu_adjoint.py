"""Hand-written adjoint of AO evaluation and of the ISDF fit (GPU).

* The AO adjoint kernel, sum_p abar[p, mu] grad_r phi_mu(r_p), is checked
  against an explicit contraction with the dx/dy/dz AO derivatives.  Those
  are the reference's gradient axes (cpp/src/ao_eval.cpp:215-244), here from
  the oracle.
* The nuclear gradient of the fitted pair-reconstruction loss is checked
  against central finite differences in FP64, h = 1e-4 Bohr, grid and anchor
  points fixed: 1e-6 relative in the max norm.
* The torch.autograd.Function returns the same gradient through backward().
* On the differentiation workload ((H2O)16, 6-31G*, level-3 grid, 2432
  anchors) the loss gradient and the Coulomb-energy gradient are checked on
  all 48 nuclei.
"""
import numpy as np
import pytest
import torch

import oracle
from mp_isdf_amd import adjoint, molecule

pytestmark = pytest.mark.gpu

H = 1e-4          # finite-difference step (Bohr)
FD_RTOL = 1e-6    # adjoint vs central differences, relative to max |gradient|


def small_case(basis="6-31g*", cart=True, npts=1500, nanchor=40, seed=5):
    mol = molecule.single_water(basis, cart=cart)
    rng = np.random.default_rng(seed)
    pts = mol.coords[rng.integers(0, mol.natm, npts)] + rng.normal(scale=1.1, size=(npts, 3))
    w = rng.uniform(0.05, 1.0, size=npts)
    anchors = np.asarray(oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", max_rank=nanchor, cart=cart))
    return mol, pts, w, anchors


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_ao_adjoint_matches_gradient_contraction(mp_isdf, basis, cart):
    mol = molecule.single_water(basis, cart=cart)
    rng = np.random.default_rng(11)
    pts = rng.normal(scale=1.5, size=(333, 3))
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=cart)
    abar = rng.normal(size=(len(pts), ctx.nao))
    got = mp_isdf.ao_adjoint(ctx, pts, abar)
    grad = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, deriv=1, cart=cart)  # [3][n][nao]
    ref = np.einsum("cpm,pm->cm", grad, abar)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_ao_adjoint_empty_and_shape_errors(mp_isdf):
    mol = molecule.single_water("sto-3g")
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    got = mp_isdf.ao_adjoint(ctx, np.zeros((0, 3)), np.zeros((0, ctx.nao)))
    assert got.shape == (3, ctx.nao) and not got.any()
    with pytest.raises(ValueError):
        mp_isdf.ao_adjoint(ctx, np.zeros((4, 3)), np.zeros((4, ctx.nao + 1)))


def test_loss_equals_squared_pair_error(mp_isdf):
    mol, pts, w, anchors = small_case()
    loss, grad, gao = adjoint.pair_loss_and_gradient(mol, pts[anchors], w[anchors], pts, w, batch_size=256)
    value = adjoint.pair_loss_value(mol, pts[anchors], w[anchors], pts, w, batch_size=512)
    np.testing.assert_allclose(loss, value, rtol=1e-10)
    assert grad.shape == (mol.natm, 3) and gao.shape == (3, mol.nao)


@pytest.mark.parametrize("basis,cart", [("6-31g*", True), ("cc-pvdz", False)])
def test_fit_gradient_matches_central_differences(mp_isdf, basis, cart):
    mol, pts, w, anchors = small_case(basis, cart)
    _, grad, _ = adjoint.pair_loss_and_gradient(mol, pts[anchors], w[anchors], pts, w, batch_size=256)
    fd = np.zeros_like(grad)
    for a in range(mol.natm):
        for c in range(3):
            vals = []
            for sgn in (1, -1):
                x = mol.coords.copy()
                x[a, c] += sgn * H
                vals.append(adjoint.pair_loss_value(adjoint.with_coords(mol, x), pts[anchors], w[anchors], pts, w))
            fd[a, c] = (vals[0] - vals[1]) / (2 * H)
    scale = np.abs(grad).max()
    assert scale > 0
    assert np.abs(grad - fd).max() <= FD_RTOL * scale, (grad, fd)


def test_autograd_function_backward(mp_isdf):
    mol, pts, w, anchors = small_case()
    coords = torch.tensor(mol.coords, dtype=torch.float64, requires_grad=True)
    loss = adjoint.pair_loss(coords, mol, pts[anchors], w[anchors], pts, w, batch_size=256)
    loss.backward()
    _, grad, _ = adjoint.pair_loss_and_gradient(mol, pts[anchors], w[anchors], pts, w, batch_size=256)
    np.testing.assert_allclose(coords.grad.numpy(), grad, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(loss.item(), adjoint.pair_loss_value(mol, pts[anchors], w[anchors], pts, w),
                               rtol=1e-10)


@pytest.fixture(scope="module")
def h2o16(mp_isdf):
    """The differentiation workload: (H2O)16, 6-31G* Cartesian (304 AOs),
    level-3 grid (551,695 kept points), 2432 anchors chosen by rand_sketch,
    and a fixed density matrix 2 B B^T (B: 304 x 40, seed 3)."""
    from mp_isdf_amd import grid as gridmod
    mol = molecule.water_cluster(16, "6-31g*", cart=True)
    assert mol.nao == 304 and mol.natm == 48
    xyz, w = gridmod.make_grid(mol, 3)
    keep = w != 0
    pts, wk = xyz[keep], w[keep]
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=True)
    sel = mp_isdf.AnchorSelection(ctx)
    k = 2432
    anchors = np.asarray(sel.select(pts, wk, mp_isdf.AxisType.AtomicOrbital,
                                    algorithm=mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao,
                                    max_rank=k, oversampling=(k + 9) // 10, seed=7))
    assert len(anchors) == k
    b = np.random.default_rng(3).normal(scale=0.3, size=(mol.nao, 40))
    return mol, pts, wk, anchors, 2.0 * b @ b.T


@pytest.mark.timeout(900)
def test_differentiation_form_h2o16_loss(h2o16):
    """dL/dR of the pair-reconstruction loss on all 48 nuclei (144
    coordinates) of the differentiation workload against finite differences.

    At this size two effects bound plain central differences at h = 1e-4:
    * the FP64 loss is a ~1e5-fold cancellation (sum |w| ||phi||^4 minus the
      fitted part), reproducible to 1e-15 but carrying rounding noise of
      ~1e-10 relative, amplified by 1/(2h);
    * the O(h^2) truncation is large on oxygen coordinates (core exponent
      ~5.5e3 Bohr^-2): 1.7e-5 of max |dL/dR| at h = 2e-4, dropping as h^2.
    Richardson-extrapolated central differences remove the h^2 term; the
    larger the steps, the lower the noise.  Measured on all 144 coordinates
    (in units of max |dL/dR|): the 8e-4/4e-4 pair agrees to 1.13e-6 max
    (median 3.7e-7), the 1.6e-3/8e-4 pair to 8.6e-7 max (median 1.9e-7), and
    the three-step h^4 extrapolation over 4e-4..1.6e-3 to 1.2e-6 max; on
    hydrogen 47/y plain central differences scatter by 4e-7 .. 6.7e-6 for
    h = 1e-4 .. 5e-5 (noise ~ 1/h).
    """
    mol, pts, wk, anchors, _ = h2o16
    _, grad, _ = adjoint.pair_loss_and_gradient(mol, pts[anchors], wk[anchors], pts, wk, batch_size=8192)
    assert grad.shape == (48, 3)
    scale = np.abs(grad).max()

    def central(a, c, h):
        vals = []
        for sgn in (1, -1):
            x = mol.coords.copy()
            x[a, c] += sgn * h
            vals.append(adjoint.pair_loss_value(adjoint.with_coords(mol, x), pts[anchors], wk[anchors], pts, wk,
                                                batch_size=8192))
        return (vals[0] - vals[1]) / (2 * h)

    fd = np.array([[(4 * central(a, c, 8e-4) - central(a, c, 1.6e-3)) / 3 for c in range(3)] for a in range(48)])
    err = np.abs(grad - fd) / scale
    print(f"\n(H2O)16 dL/dR: max |dL/dR| {scale:.6e}; adjoint vs Richardson differences, all 144 coordinates: "
          f"max {err.max():.3e}, median {np.median(err):.3e} (relative to max |dL/dR|)")
    assert err.max() <= FD_RTOL, np.unravel_index(err.argmax(), err.shape)


@pytest.mark.timeout(900)
def test_differentiation_form_h2o16_coulomb_energy(h2o16):
    """dE_J/dR on all 48 nuclei (144 coordinates) of the differentiation
    workload against plain central differences at h = 1e-4 Bohr, to 1e-6 of
    max |dE_J/dR|.  E_J carries no large cancellation, so plain differences
    suffice: measured on 0/x, 3/y, 1/z, 47/y, the h = 1e-4 differences agree
    to 3e-8 .. 2e-7 (h = 8e-4: up to 8.6e-6 on oxygen, the h^2 term)."""
    mol, pts, wk, anchors, dm = h2o16
    e, grad, _ = adjoint.coulomb_energy_and_gradient(mol, pts[anchors], pts, wk, dm, batch_size=8192)
    scale = np.abs(grad).max()
    fd = np.zeros_like(grad)
    for a in range(48):
        for c in range(3):
            vals = []
            for sgn in (1, -1):
                x = mol.coords.copy()
                x[a, c] += sgn * H
                vals.append(adjoint.coulomb_energy_value(adjoint.with_coords(mol, x), pts[anchors], pts, wk, dm,
                                                         batch_size=8192))
            fd[a, c] = (vals[0] - vals[1]) / (2 * H)
    err = np.abs(grad - fd) / scale
    print(f"\n(H2O)16 dE_J/dR: E_J {e:.10e}, max |dE_J/dR| {scale:.6e}; adjoint vs central differences "
          f"(h = 1e-4), all 144 coordinates: max {err.max():.3e}, median {np.median(err):.3e}")
    assert err.max() <= FD_RTOL, np.unravel_index(err.argmax(), err.shape)
