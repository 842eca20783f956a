"""GPU CVT anchors (cpp/src/anchor_select_cvt.cpp:17-204) on the reference's
CVT cases (python/test/test_cvt_anchor_select.py:29-74) and identical to the
oracle's on a water grid and a random cloud."""
import numpy as np
import pytest

import oracle
import cvt_cases as cases
from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


def gpu_cvt(mp_isdf):
    mol = molecule.hydrogen_molecule("sto-3g")
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))

    def select(points, weights, cvt_weights, max_rank=2, **kwargs):
        return sel.select(points, weights, mp_isdf.AxisType.AtomicOrbital, max_rank=max_rank,
                          algorithm=mp_isdf.LocalAnchorSelectionAlgorithm.CentroidalVoronoi,
                          cvt_weights=cvt_weights, **kwargs)
    return select


def test_reference_cases(mp_isdf):
    select = gpu_cvt(mp_isdf)
    cases.check_weighted_clusters_and_pruned_snap(select)
    cases.check_validation_and_degenerate_points(select)
    cases.check_exclusive_weight_modes(select)
    with pytest.raises(ValueError):
        select(np.zeros((4, 3)), np.ones(4), None)


@pytest.mark.parametrize("rank,seed", [(40, 7), (1100, 3)])
def test_water_grid_matches_oracle(mp_isdf, rank, seed):
    """Level-1 water grid (8360 points); 1100 centres take two LDS tiles in
    the assignment kernel."""
    mol = molecule.single_water("cc-pvdz")
    xyz, w = gridmod.make_grid(mol, 1)
    phi = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz)
    density = np.sum(phi ** 2, axis=1)
    select = gpu_cvt(mp_isdf)
    got = select(xyz, w, density, max_rank=rank, seed=seed, cvt_max_iter=30)
    want = oracle.select_cvt(xyz, w, density, rank, seed=seed, cvt_max_iter=30)
    assert got == want
    assert got == sorted(set(got)) and all(w[i] != 0 for i in got)


def test_random_cloud_matches_oracle(mp_isdf):
    rng = np.random.default_rng(17)
    pts = rng.normal(scale=3.0, size=(5000, 3))
    w = rng.uniform(-1, 1, 5000)
    w[::7] = 0
    m = rng.uniform(0, 2, 5000)
    select = gpu_cvt(mp_isdf)
    got = select(pts, w, m, max_rank=64, seed=5, cvt_tol=0.0, cvt_max_iter=50)
    assert got == oracle.select_cvt(pts, w, m, 64, seed=5, cvt_tol=0.0, cvt_max_iter=50)


def test_driver_weight_modes(mp_isdf):
    """test_cvt_anchor_select.py:67-74 with the quadrature convention (the
    electron density needs PySCF and raises NotImplementedError here)."""
    from mp_isdf_amd.isdf import build_isdf
    mol = molecule.hydrogen_molecule("sto-3g")
    res = build_isdf(mol, level=1, anchor_rank=2, anchor_algorithm="cvt", cvt_weight="quadrature_weights",
                     selection_only=True)
    assert len(res["anchor_indices"]) == 2 and np.all(res["anchor_weights"] != 0)
    xyz, w = gridmod.make_grid(mol, 1)
    assert list(res["anchor_indices"]) == oracle.select_cvt(xyz, w, np.abs(w), 2, seed=7)
    with pytest.raises(NotImplementedError):
        build_isdf(mol, level=1, anchor_rank=2, anchor_algorithm="cvt", selection_only=True)
    with pytest.raises(ValueError, match="positive anchor_rank"):
        build_isdf(mol, level=1, anchor_algorithm="cvt", selection_only=True)
