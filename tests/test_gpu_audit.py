"""The dense audit of sketch pivots on the GPU
(cpp/src/anchor_select_randomized_id.cpp:19-88, :164-165, :389-400): the
projection residuals of the selected and of equally many dense QRCP pivots
on the actual weighted pair matrix, against the oracle's restatement, for
the quadrature and the density policies, and the 2-million-element limit."""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


def water_points(n=600, seed=5):
    rng = np.random.default_rng(seed)
    pts = rng.normal(scale=1.5, size=(n, 3))
    w = rng.uniform(-0.2, 1.0, size=n)
    w[::50] = 0.0  # zero-weight points are not candidates
    return pts, w


def test_audit_quadrature_matches_oracle(mp_isdf):
    mol = molecule.single_water("cc-pvdz")
    pts, w = water_points()
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    alg = mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao
    anchors, diag = sel.select_with_diagnostics(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, max_rank=40,
                                                oversampling=10, seed=7, audit=True)
    plain = sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, max_rank=40, oversampling=10, seed=7)
    assert anchors == plain  # the audit does not change the selection
    kept = np.flatnonzero(w != 0)
    local = np.searchsorted(kept, anchors)
    f = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts[kept])
    actual, dense = oracle.audit_projection(f, np.sqrt(np.abs(w[kept])), local)
    print(f"\naudit rank 40: selected {diag['audit'][0]:.6e} (oracle {actual:.6e}), "
          f"dense QRCP {diag['audit'][1]:.6e} (oracle {dense:.6e})")
    np.testing.assert_allclose(diag["audit"], [actual, dense], rtol=1e-9)
    assert "audit_density" not in diag


@pytest.mark.parametrize("algorithm", ["RandomizedAODensityKhatriRao", "RandomizedAODensityNormalizedKhatriRao"])
def test_audit_density_matches_oracle(mp_isdf, algorithm):
    mol = molecule.single_water("cc-pvdz")
    pts, w = water_points(500, seed=8)
    imp = np.random.default_rng(3).uniform(0.0, 2.0, size=len(w))
    imp[::37] = 0.0
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    anchors, diag = sel.select_with_diagnostics(pts, w, mp_isdf.AxisType.AtomicOrbital,
                                                algorithm=getattr(mp_isdf.LocalAnchorSelectionAlgorithm, algorithm),
                                                max_rank=30, oversampling=8, seed=7, audit=True,
                                                point_importance=imp)
    kept = np.flatnonzero((w != 0) & (imp != 0))
    f = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts[kept])
    q = np.sum(f * f, axis=1)
    assert np.all(q > 1e-24)  # nothing screened: the retained grid is `kept`
    local = np.searchsorted(kept, anchors)
    col = np.sqrt(imp[kept]) / (q if "Normalized" in algorithm else 1.0)
    np.testing.assert_allclose(diag["audit"], oracle.audit_projection(f, np.sqrt(np.abs(w[kept])), local), rtol=1e-9)
    np.testing.assert_allclose(diag["audit_density"], oracle.audit_projection(f, col, local), rtol=1e-9)


def test_audit_limits_and_reference_cases(mp_isdf, capfd):
    mol = molecule.single_water("cc-pvdz")  # 24 AOs: 300 packed pair rows
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    alg = mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao
    pts, w = water_points(7000, seed=2)
    with pytest.raises(ValueError, match="2 million"):
        sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, max_rank=4, audit=True)
    # test_randomized_anchor_select.py:178-179: an audited rank-2 selection.
    pts, w = water_points(60, seed=4)
    got = sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, max_rank=2, audit=True)
    assert len(got) == 2
    err = capfd.readouterr().err
    assert "[randomized id audit] objective=physical_quadrature rank=2" in err
