"""CPU checks of the oracle's grid Coulomb metric (python/isdf.py:48-67,
:76-92), J/K contractions (:130-151) and the analytic s-Gaussian integrals
used as the C1 (H2 / STO-3G) J/K reference.

What pins them:
* the reference's own cross-metric fixture (test_isdf.py:205-214) and a
  hand-written 3-point kernel matrix;
* the Coulomb self-energy of a normalised Gaussian charge, sqrt(2 alpha / pi),
  on a Becke grid: a closed form that also needs the spherical-cell self
  term (its share is asserted to exceed the tolerance);
* the textbook STO-3G H2 integrals at R = 1.4 bohr (Szabo and Ostlund,
  Modern Quantum Chemistry, section 3.5.2: S12 = 0.6593, (11|11) = 0.7746,
  (11|22) = 0.5697, (12|12) = 0.2970, (11|12) = 0.4441), the one-centre
  primitive value 2 sqrt(alpha / pi) and the 1/R far-field limit;
* the J/K contraction identity of test_isdf.py:236-238.
"""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import mp_isdf as module
from mp_isdf_amd.basis import make_tables
from mp_isdf_amd.molecule import BOHR, hydrogen_molecule

C_SPHERE = (32.0 / 15.0) * np.pi ** 2 * (3.0 / (4.0 * np.pi)) ** (5.0 / 3.0)


def test_cross_metric_reference_fixture():
    # test_isdf.py:205-214
    coords = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.]])
    weights = np.array([0.3, -0.2, 0.5])
    left = np.array([[1., 2., -1.]])
    right = np.array([[0., 1., 3.], [2., 0., 1.]])
    stacked = oracle.grid_coulomb_metric(np.concatenate((left, right)), coords, weights)
    np.testing.assert_allclose(oracle.grid_cross_metric(left, right, coords, weights), stacked[:1, 1:])
    np.testing.assert_allclose(oracle.grid_cross_metric(right, left, coords, weights), stacked[1:, :1])
    # The same case by hand: the kernel matrix of the three points.
    kern = np.array([[C_SPHERE * 0.3 ** (-1 / 3), 1.0, 1.0],
                     [1.0, C_SPHERE * 0.2 ** (-1 / 3), 2 ** -0.5],
                     [1.0, 2 ** -0.5, C_SPHERE * 0.5 ** (-1 / 3)]])
    wl = left * weights
    wr = right * weights
    np.testing.assert_allclose(oracle.grid_cross_metric(left, right, coords, weights), wl @ kern @ wr.T,
                               rtol=1e-14)


def test_coincident_points_have_zero_kernel():
    coords = np.array([[0., 0., 0.], [0., 0., 0.], [0., 0., 2.]])
    weights = np.array([0.4, 0.7, 0.2])
    v = np.eye(3)
    got = oracle.grid_cross_metric(v, v, coords, weights)
    want = np.diag(C_SPHERE * np.abs(weights) ** (-1 / 3)) + np.array([[0, 0, 0.5], [0, 0, 0.5], [0.5, 0.5, 0]])
    np.testing.assert_allclose(got, weights[:, None] * want * weights[None, :], rtol=1e-14)


@pytest.mark.parametrize("alpha", [0.5, 1.0, 2.0])
def test_gaussian_self_energy_on_becke_grid(alpha):
    mol = make_tables(["H"], np.zeros((1, 3)), "sto-3g")
    xyz, w = gridmod.make_grid(mol, 3)
    rho = (alpha / np.pi) ** 1.5 * np.exp(-alpha * np.sum(xyz ** 2, axis=1))
    exact = np.sqrt(2 * alpha / np.pi)
    got = oracle.grid_coulomb_metric(rho[None, :], xyz, w)[0, 0]
    tol = 1.5e-3
    assert abs(got - exact) < tol * exact
    # Without the spherical-cell self term the result would miss by more than tol.
    self_term = np.sum((rho * w) ** 2 * C_SPHERE * np.abs(w) ** (-1 / 3))
    assert self_term > 2 * tol * exact


def test_s_integrals_match_textbook_h2():
    mol = hydrogen_molecule("sto-3g", half_bond_angstrom=0.7 * BOHR)
    funcs = oracle.s_functions(mol.atm, mol.bas, mol.env)
    s = oracle.s_overlap(funcs)
    np.testing.assert_allclose(np.diag(s), 1.0, rtol=1e-6)
    assert abs(s[0, 1] - 0.6593) < 2e-4
    eri = oracle.s_eri(funcs)
    for idx, value in (((0, 0, 0, 0), 0.7746), ((0, 0, 1, 1), 0.5697), ((0, 1, 0, 1), 0.2970),
                       ((0, 0, 0, 1), 0.4441)):
        assert abs(eri[idx] - value) < 2e-4, (idx, eri[idx])
    # 8-fold permutational symmetry
    np.testing.assert_allclose(eri, eri.transpose(1, 0, 2, 3), rtol=1e-14)
    np.testing.assert_allclose(eri, eri.transpose(2, 3, 0, 1), rtol=1e-14)


def test_s_integrals_one_centre_and_far_field():
    alpha = 0.8
    norm = (2 * alpha / np.pi) ** 0.75
    one = [(np.zeros(3), np.array([alpha]), np.array([norm]))]
    assert abs(oracle.s_eri(one)[0, 0, 0, 0] - 2 * np.sqrt(alpha / np.pi)) < 1e-14
    far = one + [(np.array([0.0, 0.0, 60.0]), np.array([alpha]), np.array([norm]))]
    np.testing.assert_allclose(oracle.s_eri(far)[0, 0, 1, 1], 1 / 60.0, rtol=1e-12)


def test_jk_contraction_identity():
    rng = np.random.default_rng(11)
    k, nao = 7, 5
    x = rng.normal(size=(k, nao))
    z = rng.normal(size=(k, k))
    z = z + z.T
    dm = rng.normal(size=(nao, nao))
    dm = dm + dm.T
    j, kk = oracle.evaluate_isdf_jk(x, z, dm)
    eri = np.einsum("pu,pv,pq,qw,qx->uvwx", x, x, z, x, x)
    np.testing.assert_allclose(j, np.einsum("uvwx,wx->uv", eri, dm), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(kk, np.einsum("uwvx,wx->uv", eri, dm), rtol=1e-12, atol=1e-12)


def h2_rhf_density(funcs):
    """Closed-form RHF density of minimal-basis H2: the bonding orbital
    (phi1 + phi2) / sqrt(2 (1 + S12)), doubly occupied."""
    s12 = oracle.s_overlap(funcs)[0, 1]
    c = np.array([1.0, 1.0]) / np.sqrt(2 * (1 + s12))
    return 2 * np.outer(c, c)


def test_c1_jk_of_the_oracle_pipeline_is_close_to_exact():
    """H2 / STO-3G, level-3 grid, qrcp_dense with 3 anchors (full pair rank):
    the ISDF J/K differ from the analytic ones only by the grid metric's
    quadrature error (1.0e-3 for J and K measured on the level-3 Lebedev
    grid; the reference's own test checks only that they are finite)."""
    mol = hydrogen_molecule("sto-3g")
    funcs = oracle.s_functions(mol.atm, mol.bas, mol.env)
    dm = h2_rhf_density(funcs)
    eri = oracle.s_eri(funcs)
    xyz, w = gridmod.make_grid(mol, 3)
    keep = w != 0
    anchors = oracle.select(mol.atm, mol.bas, mol.env, xyz, w, "ao", max_rank=3, algorithm="qrcp_dense")
    ax = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz[anchors])
    aw = w[anchors]
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, xyz[keep])
    pairs = (ax[:, :, None] * ax[:, None, :]).reshape(3, -1)
    theta, _, _ = oracle.physical_pair_refit(pairs, aw, ao, 4096)
    z = aw[:, None] * oracle.grid_coulomb_metric(theta, xyz[keep], w[keep]) * aw[None, :]
    j, k = oracle.evaluate_isdf_jk(ax, z, dm)
    dj = np.abs(j - np.einsum("uvwx,wx->uv", eri, dm)).max()
    dk = np.abs(k - np.einsum("uwvx,wx->uv", eri, dm)).max()
    print(f"\nC1 max|dJ| {dj:.3e} max|dK| {dk:.3e}")
    assert dj < 2e-3 and dk < 2e-3


@pytest.mark.skipif(module.device_enabled, reason="checks the no-device path")
def test_gpu_entry_points_fail_loudly_without_a_device():
    with pytest.raises(RuntimeError):
        module.grid_coulomb_metric(np.ones((1, 2)), np.zeros((2, 3)), np.ones(2))
    with pytest.raises(RuntimeError):
        module.isdf_jk(np.ones((1, 2)), np.ones((1, 1)), np.eye(2))
    with pytest.raises(RuntimeError):
        module.rbrp_ao(np.ones((4, 2)), np.ones(4), max_rank=1)
    mol = hydrogen_molecule("sto-3g")
    with pytest.raises(RuntimeError):
        module.LibcintContext(mol.atm, mol.bas, mol.env)
    from mp_isdf_amd import capi
    import ctypes
    out = np.zeros(2, dtype=np.int64)
    count = ctypes.c_int64()
    with pytest.raises(RuntimeError):
        capi.call("mpisdf_select_cvt", np.zeros((4, 3)), np.ones(4), np.ones(4), 4, 2, 7, 10, 1e-5, out, 2,
                  ctypes.byref(count))


def test_driver_metric_argument_validation():
    """build_isdf's metric options are checked before any device work
    (python/isdf.py:463-468)."""
    from mp_isdf_amd import isdf as driver
    mol = hydrogen_molecule("sto-3g")
    grid = (np.zeros((4, 3)), np.ones(4))
    with pytest.raises(ValueError, match="unknown metric algorithm"):
        driver.build_isdf(mol, grid=grid, compute_metric=True, metric_algorithm="bogus")
    with pytest.raises(ValueError, match="selection_only"):
        driver.build_isdf(mol, grid=grid, selection_only=True, compute_metric=True)
    with pytest.raises(ValueError, match="metric_seed"):
        driver.build_isdf(mol, grid=grid, compute_metric=True, metric_algorithm="unblocked_uniform", metric_seed=-1)
    with pytest.raises(ValueError, match="metric"):
        driver.evaluate_isdf_jk({"X1": np.ones((1, 2)), "coulomb_metric": None}, np.eye(2))


def test_driver_space_argument():
    """build_isdf(space=...) takes the reference's values; the host space is
    not part of this build (python/isdf.py:288)."""
    from mp_isdf_amd import isdf as driver
    mol = hydrogen_molecule("sto-3g")
    grid = (np.zeros((4, 3)), np.ones(4))
    with pytest.raises(ValueError, match="host"):
        driver.build_isdf(mol, grid=grid, space="host")
    with pytest.raises(ValueError, match="space"):
        driver.build_isdf(mol, grid=grid, space="gpu")
