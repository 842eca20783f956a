"""The density Khatri-Rao sketch cases of the reference's
python/test/test_density_sketch.py:15-121, restated without PySCF: the same
He / STO-3G five-point line and the H2 / STO-3G 23-point random case, with
tables from mp_isdf_amd.basis.  ``select`` is any callable with the
AnchorSelection.select keyword surface (the oracle's or the GPU module's);
``ao`` returns AO values of the case's basis on given points."""
import numpy as np
import pytest

from mp_isdf_amd.basis import make_tables
from mp_isdf_amd.molecule import BOHR

HE_POINTS = np.array([[0., 0, 0], [1., 0, 0], [2., 0, 0], [3., 0, 0], [1e3, 0, 0]])
HE_WEIGHTS = np.array([1., -2., 4., 0., 1.])
HE_MASS = np.array([1., 2., 3., 1e6, 1e6])


def helium():
    return make_tables(["He"], np.zeros((1, 3)), "sto-3g")


def hydrogen_pair():
    # gto.M(atom="H 0 0 0; H 0 0 0.74") (Angstrom)
    return make_tables(["H", "H"], np.array([[0., 0, 0], [0., 0, 0.74]]) / BOHR, "sto-3g")


def he_select(select, normalized=True, **kwargs):
    opts = dict(algorithm="density_col_norm" if normalized else "density", max_rank=1, oversampling=3,
                point_importance=HE_MASS, seed=31)
    opts.update(kwargs)
    return select(HE_POINTS, HE_WEIGHTS, **opts)


def check_rank_one_cases(select, ao):
    """test_density_sketch.py:35-61: normalized selection picks the greatest
    mass; raw selection the greatest mass * q^2; no hidden quadrature factor."""
    for batch_size in (1, 2, 5):
        assert he_select(select, batch_size=batch_size) == [2]
        raw = he_select(select, False, batch_size=batch_size)
        q = np.sum(ao(HE_POINTS) ** 2, axis=1)
        expected = int(np.argmax((HE_MASS * q ** 2) * (HE_WEIGHTS != 0)))
        assert raw == [expected] and raw != [2]
    assert he_select(select, point_importance=HE_MASS * 100) == [2]
    altered = HE_WEIGHTS.copy()
    altered[:3] = [1e9, -1e-9, 1e-20]
    assert select(HE_POINTS, altered, algorithm="density_col_norm", max_rank=1, oversampling=3,
                  point_importance=HE_MASS, seed=7) == [2]
    m = HE_MASS.copy()
    m[2] = 0
    assert he_select(select, point_importance=m) == [1]


def check_screening_and_validation(select, ao):
    """test_density_sketch.py:63-84."""
    q = np.sum(ao(HE_POINTS) ** 2, axis=1)
    assert he_select(select, ao_screening_threshold=(q[1] + q[2]) / 2) == [1]
    assert he_select(select, max_rank=0, ao_screening_threshold=1e10) == []
    assert he_select(select, max_rank=0, point_importance=np.zeros(5)) == []
    for kwargs in ({"point_importance": None}, {"point_importance": np.ones(4)},
                   {"point_importance": np.array([1., -1., 1., 1., 1.])},
                   {"point_importance": np.full(5, np.nan)}, {"point_importance": np.full(5, np.inf)},
                   {"ao_screening_threshold": -1.0}, {"ao_screening_threshold": np.nan},
                   {"ao_screening_threshold": 1e10}, {"max_rank": 2},
                   {"point_importance": np.zeros(5)}):
        with pytest.raises(ValueError):
            he_select(select, **kwargs)
    with pytest.raises(ValueError, match="axes|axis"):
        select(HE_POINTS, HE_WEIGHTS, axes=["ao", "dx"], algorithm="density_col_norm", max_rank=1, oversampling=3,
               point_importance=HE_MASS)


def check_raw_matches_quadrature(select):
    """test_density_sketch.py:86-93: the raw density sketch is the quadrature
    sketch with the importance as the weights."""
    mass = HE_MASS.copy()
    mass[HE_WEIGHTS == 0] = 0
    old = select(HE_POINTS, mass, algorithm="rand_sketch", max_rank=1, oversampling=3, seed=31)
    assert he_select(select, False) == old


def check_normalized_matches_rescaled_quadrature(select, ao):
    """test_density_sketch.py:95-121: the normalized sketch equals the
    quadrature sketch with weights mass / q^2."""
    points = np.random.default_rng(74).normal(size=(23, 3))
    q = np.sum(ao(points) ** 2, axis=1)
    mass = np.linspace(.1, 2, len(points))
    weights = np.linspace(-1, 2, len(points))
    weights[2] = 0
    mass[5] = 0
    reference_weights = mass / q ** 2
    reference_weights[weights == 0] = 0
    for seed in (7, 31):
        reference = select(points, reference_weights, algorithm="rand_sketch", max_rank=2, oversampling=4,
                           seed=seed, batch_size=3)
        for batch_size in (1, 11):
            actual = select(points, weights, algorithm="density_col_norm", max_rank=2, oversampling=4, seed=seed,
                            batch_size=batch_size, point_importance=mass)
            assert actual == reference, (seed, batch_size, actual, reference)
