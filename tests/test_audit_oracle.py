"""The oracle's restatement of the dense audit of sketch pivots
(cpp/src/anchor_select_randomized_id.cpp:19-88), pinned by what LAPACK and the
mathematics fix: the packed pair matrix has the Gram (F F^T)o(F F^T) scaled by
the column weights, its dense pivots are LAPACK dgeqp3's (through SciPy), and
the projection residual equals the orthogonal projector's from NumPy's QR."""
import numpy as np
import pytest
import scipy.linalg

import oracle


def case(seed=0, n=80, d=6):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(n, d))
    col = rng.uniform(0.2, 1.5, size=n)
    return f, col


def test_packed_pairs_have_the_full_pair_gram():
    f, col = case()
    a = oracle.packed_pairs(f, col)
    assert a.shape == (21, 80)
    want = (f @ f.T) ** 2 * np.outer(col, col)
    np.testing.assert_allclose(a.T @ a, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # A plausible mistake: packing the off-diagonal pairs without sqrt(2).
    mu, nu = np.triu_indices(6)
    wrong = (f[:, mu] * f[:, nu] * col[:, None]).T
    assert not np.allclose(wrong.T @ wrong, want, rtol=1e-6)


def projector_error(a, pivots):
    if len(pivots) == 0:
        return 1.0
    q, r = np.linalg.qr(a[:, pivots])
    keep = np.abs(np.diag(r)) > 1e-12 * np.abs(r).max()
    q = q[:, keep]
    return np.linalg.norm(a - q @ (q.T @ a)) / np.linalg.norm(a)


@pytest.mark.parametrize("k", [0, 1, 5, 12, 21])
def test_projection_error_equals_the_projector(k):
    f, col = case(1)
    a = oracle.packed_pairs(f, col)
    piv = np.random.default_rng(k).choice(80, size=k, replace=False)
    got = oracle.projection_error(a, list(piv))
    want = projector_error(a, piv)
    assert abs(got - want) <= 1e-12 + 1e-10 * want
    if k == 21:  # a full-rank pivot set spans the 21 pair rows
        assert got < 1e-13
    if k == 0:
        assert got == 1.0


def test_dependent_pivots_add_no_direction():
    f, col = case(2)
    a = oracle.packed_pairs(f, col)
    piv = [3, 17, 40]
    assert oracle.projection_error(a, piv + [17, 3]) == pytest.approx(oracle.projection_error(a, piv), rel=1e-12)
    # a mistake that keeps the dependent columns normalises round-off noise
    # into spurious directions: the drop rule is what prevents it
    b = a.copy()
    b[:, 5] = 2.0 * b[:, 3]
    assert oracle.projection_error(b, [3, 5]) == pytest.approx(oracle.projection_error(b, [3]), rel=1e-12)


def test_dense_pivots_are_lapack_dgeqp3():
    f, col = case(3, n=120, d=7)
    a = oracle.packed_pairs(f, col)
    _, _, piv = scipy.linalg.qr(a, pivoting=True, mode="economic")
    _, mine, _ = oracle.geqp3(a)
    np.testing.assert_array_equal(mine[:28], piv[:28])
    sel = [0, 1, 2, 3, 4, 5, 6, 7]
    actual, dense = oracle.audit_projection(f, col, sel)
    assert actual == pytest.approx(projector_error(a, np.array(sel)), rel=1e-10)
    assert dense == pytest.approx(projector_error(a, piv[:8]), rel=1e-10)
    # Greedy pivots beat an arbitrary set of the same size here.
    assert dense < actual


def test_driver_audit_needs_a_sketch():
    """python/isdf.py:461-462: anchor_audit is for the rand_sketch algorithms."""
    from mp_isdf_amd import isdf as driver
    from mp_isdf_amd import molecule
    mol = molecule.single_water("sto-3g")
    with pytest.raises(ValueError, match="rand sketch"):
        driver.build_isdf(mol, grid=(np.zeros((4, 3)), np.ones(4)), anchor_algorithm="qrcp_dense", anchor_audit=True)
