"""GPU density Khatri-Rao sketches (RandomizedAODensityKhatriRao and
RandomizedAODensityNormalizedKhatriRao, cpp/src/anchor_select_randomized_id.cpp:92-374)
on the reference's own density-sketch cases (python/test/test_density_sketch.py:35-121)
and against the oracle's anchors on a water grid."""
import numpy as np
import pytest

import oracle
import density_cases as cases
from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu

NAMES = {"rand_sketch": "RandomizedQuadratureKhatriRao", "density": "RandomizedAODensityKhatriRao",
         "density_col_norm": "RandomizedAODensityNormalizedKhatriRao"}
AXES = {"ao": "AtomicOrbital", "dx": "AtomicOrbitalGradX"}


def gpu_select(mp_isdf, mol):
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    sel = mp_isdf.AnchorSelection(ctx)

    def select(points, weights, axes="ao", algorithm="rand_sketch", **kwargs):
        axis = [getattr(mp_isdf.AxisType, AXES[a]) for a in axes] if isinstance(axes, list) else \
            getattr(mp_isdf.AxisType, AXES[axes])
        return sel.select(points, weights, axis,
                          algorithm=getattr(mp_isdf.LocalAnchorSelectionAlgorithm, NAMES[algorithm]), **kwargs)
    return select, (lambda pts: ctx.eval_ao(pts))


def test_rank_one_cases(mp_isdf):
    select, ao = gpu_select(mp_isdf, cases.helium())
    cases.check_rank_one_cases(select, ao)


def test_screening_and_validation(mp_isdf):
    select, ao = gpu_select(mp_isdf, cases.helium())
    cases.check_screening_and_validation(select, ao)


def test_raw_matches_quadrature(mp_isdf):
    select, _ = gpu_select(mp_isdf, cases.helium())
    cases.check_raw_matches_quadrature(select)


def test_normalized_matches_rescaled_quadrature(mp_isdf):
    select, ao = gpu_select(mp_isdf, cases.hydrogen_pair())
    cases.check_normalized_matches_rescaled_quadrature(select, ao)


@pytest.mark.parametrize("algorithm", ["density", "density_col_norm"])
def test_water_density_anchors_match_oracle(mp_isdf, algorithm):
    """(H2O) cc-pVDZ, level-1 grid, a positive importance shaped like a
    density (sum of squared AO values, zero on a tenth of the points), 60
    anchors: FP64 anchors identical to the oracle's; BF16 anchors valid."""
    mol = molecule.single_water("cc-pvdz")
    xyz, w = gridmod.make_grid(mol, 1)
    phi = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz)
    importance = np.sum(phi ** 2, axis=1) * 0.37
    importance[::10] = 0.0
    select, _ = gpu_select(mp_isdf, mol)
    opts = dict(max_rank=60, oversampling=8, seed=7, point_importance=importance)
    got = select(xyz, w, algorithm=algorithm, **opts)
    want = oracle.select(mol.atm, mol.bas, mol.env, xyz, w, "ao", algorithm=algorithm, **opts)
    assert got == want
    low = select(xyz, w, algorithm=algorithm, precision="bf16", **opts)
    assert len(low) == 60 and len(set(low)) == 60
    assert all(w[i] != 0 and importance[i] > 0 for i in low)
    print(f"{algorithm}: bf16 anchors shared with fp64: {len(set(low) & set(got))}/60")


def test_driver_density_sketch(mp_isdf):
    """build_isdf(anchor_algorithm="rand_sketch", sketch="density_col_norm")
    with a given importance selects what the module does, and the
    quadrature_weights convention equals the raw density sketch on |w|."""
    from mp_isdf_amd.isdf import build_isdf
    mol = molecule.single_water("sto-3g")
    xyz, w = gridmod.make_grid(mol, 1)
    importance = np.abs(w) * np.linspace(0.5, 1.5, len(w))
    res = build_isdf(mol, grid=(xyz, w), anchor_algorithm="rand_sketch", sketch="density_col_norm", anchor_rank=12,
                     anchor_oversampling=4, point_importance=importance, selection_only=True)
    select, _ = gpu_select(mp_isdf, mol)
    want = select(xyz, w, algorithm="density_col_norm", max_rank=12, oversampling=4, seed=7,
                  point_importance=importance)
    np.testing.assert_array_equal(res["anchor_indices"], want)
    res = build_isdf(mol, grid=(xyz, w), anchor_algorithm="rand_sketch", sketch="density", anchor_rank=12,
                     anchor_oversampling=4, density_weight="quadrature_weights", selection_only=True)
    want = select(xyz, w, algorithm="density", max_rank=12, oversampling=4, seed=7, point_importance=np.abs(w))
    np.testing.assert_array_equal(res["anchor_indices"], want)
    with pytest.raises(NotImplementedError, match="PySCF"):
        build_isdf(mol, grid=(xyz, w), anchor_algorithm="rand_sketch", sketch="density", anchor_rank=12,
                   selection_only=True)
    with pytest.raises(ValueError, match="point_importance"):
        build_isdf(mol, grid=(xyz, w), anchor_algorithm="rand_sketch", sketch="density", anchor_rank=12,
                   point_importance=-importance, selection_only=True)
