"""CPU: the oracle's CVT selection (cpp/src/anchor_select_cvt.cpp:17-204) on the
reference's own CVT cases (python/test/test_cvt_anchor_select.py:29-74), and
the invariants the algorithm fixes: distinct sorted positive-mass sites, and
a converged Lloyd energy no worse than the seeding's on a random cloud."""
import numpy as np

import oracle
import cvt_cases as cases


def oracle_cvt(points, weights, cvt_weights, max_rank=2, **kwargs):
    return oracle.select_cvt(points, weights, cvt_weights, max_rank, **kwargs)


def test_weighted_clusters_and_pruned_snap():
    cases.check_weighted_clusters_and_pruned_snap(oracle_cvt)


def test_validation_and_degenerate_points():
    cases.check_validation_and_degenerate_points(oracle_cvt)


def test_exclusive_weight_modes():
    cases.check_exclusive_weight_modes(oracle_cvt)


def test_three_separated_clusters_get_one_site_each():
    rng = np.random.default_rng(5)
    centres = np.array([[0., 0, 0], [20., 0, 0], [0., 20, 0]])
    points = np.concatenate([c + rng.normal(scale=0.5, size=(40, 3)) for c in centres])
    w = np.ones(len(points))
    sel = oracle_cvt(points, w, np.ones(len(points)), max_rank=3, seed=11)
    assert sel == sorted(sel) and len(set(sel)) == 3
    assert sorted(i // 40 for i in sel) == [0, 1, 2]
