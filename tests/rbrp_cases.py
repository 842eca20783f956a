"""RBRP checks of the reference's python/test/test_rbrp_id.py:17-59, for any
``rbrp(ao, weights, **opts)`` with the rbrp_ao_cpu keyword surface and result
dict (the oracle's or the GPU module's)."""
import numpy as np


def check_direct_fit_at_pivots(rbrp):
    """test_rbrp_id.py:42-59: at its own pivots the ID equals the direct
    pseudo-inverse fit; the relative residual is the fit's."""
    rng = np.random.default_rng(52)
    ao = rng.normal(size=(36, 5))
    weights = np.exp(rng.normal(size=36)) * rng.choice([-1., 1.], size=36)
    opts = dict(block_size=4, batch_size=8, max_rank=8, seed=12, tolerance=0., compute_id=True)
    result = rbrp(ao, weights, **opts)
    j = np.asarray(result["indices"])
    assert len(np.unique(j)) == len(j) == 8
    np.testing.assert_array_equal(j, rbrp(ao, weights, **opts)["indices"])
    pairs = (ao[:, :, None] * ao[:, None, :]).reshape(len(ao), -1)
    direct = np.linalg.pinv(weights[j, None] * pairs[j], rcond=1e-12).T
    np.testing.assert_allclose(result["transform"], direct, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(result["interp_vecs"], (pairs @ direct.T).T, rtol=1e-10, atol=1e-10)
    residual = pairs - result["interp_vecs"].T @ (weights[j, None] * pairs[j])
    relative = (np.linalg.norm(residual * np.sqrt(np.abs(weights))[:, None]) /
                np.linalg.norm(pairs * np.sqrt(np.abs(weights))[:, None]))
    assert abs(relative - result["relative_residual"]) < 1e-10
    assert result["retained_rank"] == len(j)
    assert result["history"].shape[1] == 4
    return result


def check_absolute_weighted_selection(rbrp):
    """test_rbrp_id.py:17-40: selection depends on |w| only; the ID is
    cardinal at the anchors (theta[:, j] w_j = I)."""
    rng = np.random.default_rng(19)
    ao = rng.normal(size=(40, 4))
    w = np.exp(rng.normal(size=40)) * rng.choice([-1., 1.], size=40)
    res = rbrp(ao, w, block_size=4, max_rank=6, tolerance=0., seed=3, compute_id=True)
    other = rbrp(ao, np.abs(w), block_size=4, max_rank=6, tolerance=0., seed=3, batch_size=11)
    np.testing.assert_array_equal(res["indices"], other["indices"])
    j = np.asarray(res["indices"])
    np.testing.assert_allclose(res["interp_vecs"][:, j] * w[j][None, :], np.eye(len(j)), atol=1e-12)
    pairs = (ao[:, :, None] * ao[:, None, :]).reshape(40, -1)
    err = np.linalg.norm((pairs - res["interp_vecs"].T @ (w[j, None] * pairs[j])) * np.sqrt(np.abs(w))[:, None])
    norm = np.sqrt(np.sum(np.abs(w) * np.sum(ao * ao, axis=1) ** 2))
    assert abs(err / norm - res["relative_residual"]) < 1e-12
    assert res["retained_rank"] == len(j)


def check_tolerance_stop(rbrp):
    """The tolerance rule (anchor_select_rbrp.cpp:235-237, :308-310): the
    full pair rank reached, the relative residual falls to roundoff."""
    rng = np.random.default_rng(7)
    ao = rng.normal(size=(60, 3))
    w = rng.uniform(0.5, 1.5, 60)
    res = rbrp(ao, w, block_size=4, tolerance=1e-8, seed=1)
    assert res["status"] == "tolerance" and len(res["indices"]) == 6
    assert res["relative_residual"] <= 1e-8
