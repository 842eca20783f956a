"""CPU: the oracle's RBRP and ID (cpp/src/anchor_select_rbrp.cpp:135-318) on
the reference's own RBRP checks (python/test/test_rbrp_id.py:17-59)."""
import oracle
import rbrp_cases as cases


def test_direct_fit_at_pivots():
    cases.check_direct_fit_at_pivots(oracle.rbrp_ao)


def test_absolute_weighted_selection():
    cases.check_absolute_weighted_selection(oracle.rbrp_ao)


def test_tolerance_stop():
    cases.check_tolerance_stop(oracle.rbrp_ao)


def test_device_branch_selects_what_the_host_branch_selects():
    """The reference's two RBRP branches (anchor_select_rbrp.cpp:186-318 on the
    host, anchor_select_rbrp_device.cpp:327-491 on a GPU build) draw one uniform
    per sample: discrete_distribution takes the first active index whose
    normalised cumulative probability reaches it, the device sampler the first
    whose prefix sum exceeds it times the total; both sample without
    replacement within a block.  Their bases differ by a rotation of the same
    span, which leaves the residuals and the ID unchanged.  Restated, they pick
    the same anchors and the same interpolation vectors."""
    import numpy as np
    rng = np.random.default_rng(52)
    ao = rng.normal(size=(300, 6))
    w = np.exp(rng.normal(size=300)) * rng.choice([-1., 1.], size=300)
    for opts in (dict(block_size=4, batch_size=8, max_rank=18, seed=12, tolerance=0., compute_id=True),
                 dict(block_size=5, seed=3, tolerance=1e-6, compute_id=True),
                 dict(block_size=16, max_rank=21, seed=7, tolerance=0., refresh_every=2, compute_id=True)):
        host = oracle.rbrp_ao(ao, w, **opts)
        dev = oracle.rbrp_ao(ao, w, branch="device", **opts)
        assert host["indices"] == dev["indices"] and host["status"] == dev["status"]
        np.testing.assert_allclose(dev["interp_vecs"], host["interp_vecs"], rtol=1e-10, atol=1e-12)
        # history[:, 3] is the downdated residual ratio: near the tolerance it
        # carries the cancellation of d - decrease (~1e-9 absolute)
        np.testing.assert_allclose(dev["history"], host["history"], rtol=1e-8, atol=1e-8)
