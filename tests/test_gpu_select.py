"""GPU anchor selection against the oracle and against scipy's dgeqp3 (the
reference tests' own oracle).  Pivot lists are compared exactly."""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import basis, molecule

pytestmark = pytest.mark.gpu


class H2:
    mol = molecule.hydrogen_molecule("sto-3g", half_bond_angstrom=0.5 * 0.529177210920)
    points = np.random.default_rng(12).normal(size=(8, 3))
    weights = np.array([1., 2., 0., -1., 1., 3., 1., 1.])


@pytest.fixture(scope="module")
def h2(mp_isdf):
    ctx = mp_isdf.LibcintContext(H2.mol.atm, H2.mol.bas, H2.mol.env)
    return ctx, mp_isdf.AnchorSelection(ctx)


AXES = {"ao": "AtomicOrbital", "dx": "AtomicOrbitalGradX", "dy": "AtomicOrbitalGradY",
        "dz": "AtomicOrbitalGradZ", "occ": "OccupiedOrbital"}


def axis_list(mp_isdf, names):
    return [getattr(mp_isdf.AxisType, AXES[n]) for n in names]


@pytest.mark.parametrize("axes,max_rank", [(["ao"], 0), (["dx", "dz"], 4), (["dz", "ao", "dz"], 3),
                                           (["ao", "dx"], 3)])
def test_dense_matches_oracle(mp_isdf, h2, axes, max_rank):
    _, sel = h2
    got = sel.select(H2.points, H2.weights, axis_list(mp_isdf, axes), max_rank=max_rank)
    want = oracle.select(H2.mol.atm, H2.mol.bas, H2.mol.env, H2.points, H2.weights, axes, max_rank=max_rank)
    assert got == want
    assert 2 not in got


def test_dense_occupied_axis_weights_and_errors(mp_isdf, h2):
    _, sel = h2
    mo = np.array([[0.55, 1.2], [0.55, -1.2]])
    axes = axis_list(mp_isdf, ["ao", "occ"])
    got = sel.select(H2.points, H2.weights, axes, max_rank=3, mo_coeff=mo, nocc=1, axis_weights=[1., 2.])
    want = oracle.select(H2.mol.atm, H2.mol.bas, H2.mol.env, H2.points, H2.weights, ["ao", "occ"], max_rank=3,
                         mo_coeff=mo, nocc=1, axis_weights=[1., 2.])
    assert got == want
    assert sel.select(H2.points, np.zeros(8), axes[:1]) == []
    for factors in ([1.], [0., 1.], [float("nan"), 1.]):
        with pytest.raises(ValueError):
            sel.select(H2.points, H2.weights, axis_list(mp_isdf, ["ao", "dx"]), axis_weights=factors)
    with pytest.raises(ValueError):
        sel.select(H2.points, H2.weights, mp_isdf.AxisType.OccupiedOrbital)
    with pytest.raises(ValueError):
        sel.select(H2.points, H2.weights, axes[:1], space="host")
    with pytest.raises(RuntimeError):
        mp_isdf.DistributedAnchorSelection(h2[0]).select(H2.points, H2.weights, axes[:1])


def test_evaluate_axis_matches_oracle(mp_isdf, h2):
    _, sel = h2
    mo = np.array([[0.55, 1.2], [0.55, -1.2]])
    names = ["dz", "occ", "ao", "dz"]
    got = sel.evaluate_axis(H2.points, axis_list(mp_isdf, names), mo_coeff=mo, nocc=1)
    want = oracle.axis_eval(H2.mol.atm, H2.mol.bas, H2.mol.env, H2.points, names, mo_coeff=mo, nocc=1)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12)


class ThreeS:
    atm, bas, env = basis.three_s_shell_tables()
    points = np.random.default_rng(23).normal(size=(19, 3))
    weights = np.linspace(.4, 1.5, 19)
    weights[2] = 0
    weights[4] = -1


def test_randomized_matches_oracle_and_rules(mp_isdf):
    c = ThreeS
    ctx = mp_isdf.LibcintContext(c.atm, c.bas, c.env)
    sel = mp_isdf.AnchorSelection(ctx)
    alg = mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao
    kw = dict(algorithm=alg, sketch_size=12, oversampling=3, seed=31)
    okw = dict(algorithm="rand_sketch", sketch_size=12, oversampling=3, seed=31)
    for batch in (1, 4, 19):
        got = sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, batch_size=batch, max_rank=3, **kw)
        assert got == oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", max_rank=3, **okw)
        assert 2 not in got and len(set(got)) == 3
    full = sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, batch_size=4, **kw)
    assert full == oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", **okw)
    assert len(full) == 6
    assert sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, tolerance=1, **kw) == []
    assert len(sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, tolerance=1, max_rank=2, **kw)) == 2
    with pytest.raises(ValueError):
        sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, max_rank=7, **kw)
    with pytest.raises(ValueError):
        sel.select(c.points, c.weights, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, sketch_size=12,
                   oversampling=0)


def test_randomized_water_matches_oracle(mp_isdf):
    mol = molecule.single_water("cc-pvdz")
    rng = np.random.default_rng(5)
    pts = rng.normal(scale=1.5, size=(3000, 3))
    w = rng.uniform(0.05, 1.0, size=3000)
    w[::17] = 0
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    sel = mp_isdf.AnchorSelection(ctx)
    alg = mp_isdf.LocalAnchorSelectionAlgorithm.RandomizedQuadratureKhatriRao
    got = sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=alg, max_rank=120, oversampling=12, seed=7)
    want = oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", algorithm="rand_sketch", max_rank=120,
                         oversampling=12, seed=7)
    assert got == want


class ThreeS17:
    atm, bas, env = basis.three_s_shell_tables()
    points = np.random.default_rng(23).normal(size=(17, 3))
    weights = np.linspace(0.5, 2, 17)
    weights[2] = 0
    weights[4] *= -1


def test_online_matches_oracle(mp_isdf):
    """python/test/test_online_anchor_select.py:44-101 cases, GPU vs oracle."""
    c = ThreeS17
    ctx = mp_isdf.LibcintContext(c.atm, c.bas, c.env)
    sel = mp_isdf.AnchorSelection(ctx)
    online = mp_isdf.LocalAnchorSelectionAlgorithm.OnlineQRColumnPivoting
    ao = mp_isdf.AxisType.AtomicOrbital
    for batch in (1, 3, 17):
        got = sel.select(c.points, c.weights, ao, algorithm=online, batch_size=batch)
        assert got == oracle.select(c.atm, c.bas, c.env, c.points, c.weights, "ao", algorithm="qrcp_online",
                                    batch_size=batch)
        assert len(got) == 6
        assert sel.select(c.points, c.weights, ao, algorithm=online, batch_size=batch, max_rank=3) == got[:3]
    axes = [ao, mp_isdf.AxisType.AtomicOrbitalGradX, mp_isdf.AxisType.OccupiedOrbital]
    opts = dict(mo_coeff=np.eye(3), nocc=1, max_rank=3)
    assert (sel.select(c.points, c.weights, axes, algorithm=online, batch_size=2, **opts)
            == sel.select(c.points, c.weights, axes, **opts))
    assert sel.select(c.points, c.weights, ao, algorithm=online, tolerance=1) == []
    for cutoff in (0.1, 0.5):
        assert (sel.select(c.points, c.weights, ao, algorithm=online, tolerance=cutoff, batch_size=2)
                == sel.select(c.points, c.weights, ao, tolerance=cutoff))
    assert sel.select(c.points, np.zeros_like(c.weights), ao, algorithm=online) == []
    with pytest.raises(ValueError):
        sel.select(c.points, c.weights, ao, algorithm=online, batch_size=0)


def test_online_water_matches_oracle(mp_isdf):
    mol = molecule.single_water("cc-pvdz")
    rng = np.random.default_rng(8)
    pts = rng.normal(scale=1.5, size=(1500, 3))
    w = rng.uniform(0.05, 1.0, size=1500)
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    sel = mp_isdf.AnchorSelection(ctx)
    online = mp_isdf.LocalAnchorSelectionAlgorithm.OnlineQRColumnPivoting
    got = sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, algorithm=online, max_rank=40, batch_size=256)
    want = oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", algorithm="qrcp_online", max_rank=40,
                         batch_size=256)
    assert got == want
    dense = sel.select(pts, w, mp_isdf.AxisType.AtomicOrbital, max_rank=40)
    assert dense == oracle.select(mol.atm, mol.bas, mol.env, pts, w, "ao", max_rank=40)


def test_online_water8_ccpvdz_bit_identical(mp_isdf):
    """The (H2O)8 / cc-pVDZ (192 AOs) level-3 grid workload with qrcp_online in
    16k-point batches and rank 8: anchors bit-identical to the oracle's
    restatement of anchor_select_online_qr.cpp (precomputed by
    tests/golden/make_online_water8.py; the grid checksum must match)."""
    import json
    from pathlib import Path
    from mp_isdf_amd import grid as gridmod
    fixture = json.loads((Path(__file__).resolve().parent / "golden" / "online_water8.json").read_text())
    mol = molecule.water_cluster(8, "cc-pvdz")
    assert mol.nao == 192
    xyz, w = gridmod.make_grid(mol, 3)
    assert len(w) == fixture["grid_points"]
    np.testing.assert_allclose([np.sum(xyz), np.sum(w)], [fixture["grid_sum_xyz"], fixture["grid_sum_w"]],
                               rtol=1e-12)
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env))
    got = sel.select(xyz, w, mp_isdf.AxisType.AtomicOrbital, max_rank=8, batch_size=16384,
                     algorithm=mp_isdf.LocalAnchorSelectionAlgorithm.OnlineQRColumnPivoting)
    assert got == fixture["anchors"]
