"""GPU RBRP and ID (cpp/src/anchor_select_rbrp.cpp:135-318,
anchor_select_rbrp_device.cpp:327-491) on the reference's RBRP checks
(python/test/test_rbrp_id.py:17-59) and against the oracle's selections."""
import numpy as np
import pytest

import oracle
import rbrp_cases as cases
from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


def test_reference_checks(mp_isdf):
    got = cases.check_direct_fit_at_pivots(mp_isdf.rbrp_ao)
    cases.check_absolute_weighted_selection(mp_isdf.rbrp_ao)
    cases.check_tolerance_stop(mp_isdf.rbrp_ao)
    rng = np.random.default_rng(52)
    ao = rng.normal(size=(36, 5))
    weights = np.exp(rng.normal(size=36)) * rng.choice([-1., 1.], size=36)
    want = oracle.rbrp_ao(ao, weights, block_size=4, batch_size=8, max_rank=8, seed=12, tolerance=0.,
                          compute_id=True)
    assert got["indices"] == want["indices"] and got["status"] == want["status"]
    np.testing.assert_allclose(got["interp_vecs"], want["interp_vecs"], rtol=1e-9, atol=1e-9)


def test_device_points_match_oracle(mp_isdf):
    """Water cc-pVDZ on a level-1 grid (nonzero-weight points, 24 AOs, 300
    pair rows), 40 anchors in blocks of 16 with refreshes."""
    mol = molecule.single_water("cc-pvdz")
    xyz, w = gridmod.make_grid(mol, 1)
    keep = w != 0
    xyz, w = xyz[keep], w[keep]
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    got = mp_isdf.rbrp_ao_device(sel, xyz, w, max_rank=40, seed=7, tolerance=0.0, compute_id=True)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz)
    want = oracle.rbrp_ao(ao, w, max_rank=40, seed=7, tolerance=0.0, compute_id=True)
    assert got["indices"] == want["indices"]
    assert got["status"] == want["status"] and got["retained_rank"] == want["retained_rank"]
    assert abs(got["relative_residual"] - want["relative_residual"]) < 1e-10
    np.testing.assert_allclose(got["history"], want["history"], rtol=1e-9, atol=1e-12)
    scale = np.abs(want["interp_vecs"]).max()
    assert np.abs(got["interp_vecs"] - want["interp_vecs"]).max() < 1e-8 * scale
    # The reference driver calls the device branch on a GPU build (isdf.py:572-576):
    # its sampler and basis, restated, select the same anchors.
    dev = oracle.rbrp_ao(ao, w, max_rank=40, seed=7, tolerance=0.0, compute_id=True, branch="device")
    assert got["indices"] == dev["indices"] and got["status"] == dev["status"]
    assert np.abs(got["interp_vecs"] - dev["interp_vecs"]).max() < 1e-8 * scale
    with pytest.raises(ValueError, match="nonzero"):
        mp_isdf.rbrp_ao_device(sel, xyz, np.where(np.arange(len(w)) == 3, 0.0, w), max_rank=4)


def test_driver_rbrp_modes(mp_isdf):
    """build_isdf(anchor_algorithm="rbrp" | "rbrp_id") (python/isdf.py:564-620):
    rbrp selects with rbrp_ao_device on the kept points and refits; rbrp_id
    keeps RBRP's ID, whose pair-space error is its relative residual and is
    no smaller than the least-squares refit's at the same anchors."""
    from mp_isdf_amd.isdf import build_isdf
    mol = molecule.single_water("sto-3g")
    xyz, w = gridmod.make_grid(mol, 1)
    keep = w != 0
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    direct = mp_isdf.rbrp_ao_device(sel, xyz[keep], w[keep], max_rank=20, seed=7)
    res = build_isdf(mol, level=1, anchor_algorithm="rbrp", anchor_rank=20, pair_space_error=True)
    np.testing.assert_array_equal(res["anchor_indices"], np.flatnonzero(keep)[direct["indices"]])
    rid = build_isdf(mol, level=1, anchor_algorithm="rbrp_id", anchor_rank=20, pair_space_error=True,
                     compute_metric=True)
    np.testing.assert_array_equal(rid["anchor_indices"], res["anchor_indices"])
    e_id = rid["pair_space_errors"]["scalar"]["relative"]
    e_fit = res["pair_space_errors"]["scalar"]["relative"]
    assert e_fit <= e_id * (1 + 1e-9)
    theta = rid["interp_vecs"]
    assert theta.shape == (20, int(keep.sum()))
    assert rid["coulomb_metric"].shape == (20, 20)
    print(f"rbrp (H2O/STO-3G, 20 anchors): ID pair error {e_id:.4e}, refit {e_fit:.4e}")
