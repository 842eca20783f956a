"""The fit's grid pass with Z = L^-1 S on INT8 digit planes (form 1 of
mpisdf_isdf_fit_device_form) against the FP64 MFMA pass (form 0) and the
oracle's restatement of the reference refit and pair error
(python/isdf.py:238-263, :176-191).

Form 1 holds each column s_g of S = (Phi_a phi_g)^2 and each row of L^-1 as
five base-254 digits per entry (within ~4.8e-13 of the largest entry of
its column or row) and sums the 15 digit-pair products exactly in INT32.
The digit rounding enters the squared error
e^2 = 1 - sum_g ||z_g||^2 / sum_g ||p_g||^4 amplified by the conditioning of
the anchor-pair Gram (z = L^-1 s cancels): with four S planes (2.4e-10) it
moved e^2 by 1.2e-10 on the water cases and by 5.2e-9 on (H2O)8 with 1,536
anchors (cond(G2) 3e9); five planes leave 1e-10 on e^2 as the bound, far
below the 1e-4 on e that defines equal fit error (e moves by the e^2 change
over 2e).  A pseudo-inverse fit keeps FP64 products."""
import ctypes

import numpy as np
import pytest

import oracle
from mp_isdf_amd import capi, molecule

pytestmark = pytest.mark.gpu


def fit_form(cctx, anchor_xyz, anchor_w, xyz, w, form, batch):
    import torch
    k, n = len(anchor_w), len(w)
    res = np.zeros(12)
    tsum = np.zeros(k)
    capi.call("mpisdf_isdf_fit_device_form", cctx.handle, torch.tensor(anchor_xyz, device="cuda"),
              np.ascontiguousarray(anchor_w), k, torch.tensor(xyz, device="cuda"), torch.tensor(w, device="cuda"),
              n, batch, ctypes.c_int32(form), tsum, res, None)
    return res, tsum


def water_case(nanchor, seed=3):
    mol = molecule.single_water("sto-3g")
    rng = np.random.default_rng(seed)
    pts = rng.normal(scale=1.4, size=(900, 3))
    w = rng.uniform(0.05, 1.0, size=900)
    w[5] = -0.3
    anchors = np.asarray(oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", max_rank=nanchor))
    return mol, pts, w, anchors


@pytest.mark.parametrize("nanchor,batch", [(6, 128), (20, 7), (20, 100000)])
def test_digit_fit_matches_fp64_and_oracle(mp_isdf, nanchor, batch):
    mol, pts, w, anchors = water_case(nanchor)
    cctx = capi.Context(mol.atm, mol.bas, mol.env)
    r0, t0 = fit_form(cctx, pts[anchors], w[anchors], pts, w, 0, batch)
    r1, t1 = fit_form(cctx, pts[anchors], w[anchors], pts, w, 1, batch)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    aa = ao[anchors]
    pairs = (aa[:, :, None] * aa[:, None, :]).reshape(len(anchors), -1)
    theta, _, _ = oracle.physical_pair_refit(pairs, w[anchors], ao, 256)
    err = oracle.stream_pair_space_error_norms(ao, ao, pairs, w[anchors], theta, w)
    print(f"\nk {nanchor} batch {batch}: relative pair error fp64 {r0[1]:.12e} digits {r1[1]:.12e} "
          f"oracle {err['relative']:.12e}")
    assert abs(r1[1] ** 2 - r0[1] ** 2) <= 1e-10
    np.testing.assert_allclose(r1[1], err["relative"], rtol=1e-6)
    total = (r0[0] / r0[1]) ** 2  # sum_g |w_g| ||phi_g||^4
    assert abs(r1[0] ** 2 - r0[0] ** 2) <= 1e-10 * total
    np.testing.assert_allclose(t1, t0, rtol=1e-7, atol=1e-9 * np.abs(t0).max())
    np.testing.assert_allclose(t1, theta @ w, rtol=1e-7, atol=1e-9 * np.abs(theta @ w).max())
    assert r1[11] == (nanchor + 127) // 128 * ((len(w) + batch - 1) // batch)  # one fold launch per row group
    assert r1[2] == r0[2] and r1[8] == r0[8] == nanchor and r1[9] == 0


def test_digit_fit_pseudo_inverse(mp_isdf):
    """Dependent anchors (40 > 28 independent pairs of water STO-3G): the rows
    of W = diag(lambda)^-1/2 V^T replace L^-1, with eigenvalues down to 1e-14
    of the largest, whose inverse square roots would amplify the digits'
    rounding to ~1e-6 of z; form 1 keeps the FP64 products there."""
    mol = molecule.single_water("sto-3g")
    rng = np.random.default_rng(3)
    pts = rng.normal(scale=1.4, size=(900, 3))
    w = rng.uniform(0.05, 1.0, size=900)
    anchors = rng.choice(900, size=40, replace=False)
    cctx = capi.Context(mol.atm, mol.bas, mol.env)
    r0, t0 = fit_form(cctx, pts[anchors], w[anchors], pts, w, 0, 256)
    r1, t1 = fit_form(cctx, pts[anchors], w[anchors], pts, w, 1, 256)
    assert r0[9] == r1[9] == 1 and r1[8] == r0[8] < 40
    print(f"\npseudo-inverse: relative pair error fp64 {r0[1]:.12e} form 1 {r1[1]:.12e}")
    np.testing.assert_array_equal(r1[:2], r0[:2])
    np.testing.assert_array_equal(t1, t0)
    assert r1[11] == r0[11] == 4  # FP64 Z products, one per batch


def test_digit_fit_h2o8_1536_anchors(mp_isdf):
    """(H2O)8 / cc-pVDZ (192 AOs), level-2 grid, 1,536 rand_sketch anchors
    (cond(G2) ~1e9): twelve row groups of L^-1 with their triangular K ranges,
    several grid batches."""
    from mp_isdf_amd import grid as gridmod
    mol = molecule.water_cluster(8, "cc-pvdz")
    xyz, w = gridmod.make_grid(mol, 2)
    keep = w != 0
    pts, wk = xyz[keep], w[keep]
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    sel = mp_isdf.AnchorSelection(ctx)
    k = 1536
    anchors = np.asarray(sel.select(pts, wk, mp_isdf.AxisType.AtomicOrbital,
                                    algorithm=mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao,
                                    max_rank=k, oversampling=(k + 9) // 10, seed=7))
    cctx = capi.Context(mol.atm, mol.bas, mol.env)
    r0, t0 = fit_form(cctx, pts[anchors], wk[anchors], pts, wk, 0, 16384)
    r1, t1 = fit_form(cctx, pts[anchors], wk[anchors], pts, wk, 1, 16384)
    rel_t = np.linalg.norm(t1 - t0) / np.linalg.norm(t0)
    print(f"\n(H2O)8 k {k} n {len(wk)} cond(G2) {r0[2]:.3e}: relative pair error fp64 {r0[1]:.12e} "
          f"digits {r1[1]:.12e} (difference {r1[1] - r0[1]:+.2e}, of e^2 {r1[1] ** 2 - r0[1] ** 2:+.2e}); "
          f"theta.w {rel_t:.2e}; "
          f"Z product {r0[10]:.4f} s FP64, {r1[10]:.4f} s INT8 digits")
    assert abs(r1[1] ** 2 - r0[1] ** 2) <= 1e-10
    assert rel_t <= 1e-6
