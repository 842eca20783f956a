"""The CVT cases of the reference's python/test/test_cvt_anchor_select.py:29-74.
``select(points, weights, cvt_weights, max_rank=2, **kwargs)`` is the oracle's
or the GPU module's CentroidalVoronoi selection."""
import numpy as np
import pytest


def check_weighted_clusters_and_pruned_snap(select):
    points = np.array([[0., 0, 0], [0.1, 0, 0], [99., 0, 0], [10., 0, 0], [10.1, 0, 0]])
    weights = np.array([1., 1., 0., -1., 1.])
    density = np.ones(5)
    result = select(points, weights, density, seed=3)
    assert len(result) == 2
    assert result[0] in (0, 1) and result[1] in (3, 4)
    assert result == select(points, weights, density, seed=3)
    assert 2 not in result


def check_validation_and_degenerate_points(select):
    points = np.zeros((4, 3))
    weights = np.array([1., 0., 2., 3.])
    assert len(set(select(points, weights, np.ones(4), max_rank=3))) == 3
    for density, rank in ((np.ones(3), 2), (np.array([1., -1., 1., 1.]), 2), (np.ones(4), 4)):
        with pytest.raises(ValueError):
            select(points, weights, density, max_rank=rank)
    with pytest.raises(ValueError):
        select(points, weights, np.ones(4), cvt_max_iter=0)
    with pytest.raises(ValueError):
        select(points, weights, np.ones(4), cvt_tol=-1)
    with pytest.raises(ValueError):
        select(points, weights, np.ones(4), max_rank=0)


def check_exclusive_weight_modes(select):
    points = np.array([[0., 0, 0], [5., 0, 0], [10., 0, 0], [20., 0, 0]])
    weights = np.array([1., 1., 100., 0.])
    density = np.array([100., 1., 1., 1000000.])
    assert select(points, weights, density, max_rank=1, seed=2) == [0]
    assert select(points, weights, np.abs(weights), max_rank=1, seed=2) == [2]
