"""GPU grid Coulomb metric (python/isdf.py:48-67, :76-92) and J/K
(:130-151) against the oracle, and the C1 case (H2 / STO-3G, qrcp_dense,
3 anchors) against the analytic integrals.

Tolerances: the metric sums n^2 FP64 products in a different order than the
oracle (MFMA partial sums per 4-wide K step, row blocks), so it is compared
at 1e-11 of the largest entry; the J/K contractions at 1e-12."""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import isdf as driver
from mp_isdf_amd import molecule
from mp_isdf_amd.molecule import hydrogen_molecule

pytestmark = pytest.mark.gpu


def random_case(n, kl, kr, seed):
    rng = np.random.default_rng(seed)
    xyz = rng.normal(scale=2.0, size=(n, 3))
    xyz[7] = xyz[3]  # two distinct grid points at one position: kernel 0
    w = rng.uniform(0.01, 1.0, size=n)
    w[::11] *= -1.0  # negative quadrature weights
    left = rng.normal(size=(kl, n))
    right = rng.normal(size=(kr, n))
    return xyz, w, left, right


def close(got, want, rel):
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= rel * scale, np.abs(got - want).max() / scale


def test_metric_reference_fixture(mp_isdf):
    # test_isdf.py:205-214
    coords = np.array([[0., 0., 0.], [1., 0., 0.], [0., 1., 0.]])
    weights = np.array([0.3, -0.2, 0.5])
    left = np.array([[1., 2., -1.]])
    right = np.array([[0., 1., 3.], [2., 0., 1.]])
    stacked = mp_isdf.grid_coulomb_metric(np.concatenate((left, right)), coords, weights)
    close(stacked, oracle.grid_coulomb_metric(np.concatenate((left, right)), coords, weights), 1e-14)
    close(mp_isdf.grid_cross_metric(left, right, coords, weights), stacked[:1, 1:], 1e-14)
    close(mp_isdf.grid_cross_metric(right, left, coords, weights), stacked[1:, :1], 1e-14)


@pytest.mark.parametrize("n,k", [(3001, 37), (1024, 130)])
def test_symmetric_metric_matches_oracle(mp_isdf, n, k):
    xyz, w, theta, _ = random_case(n, k, 1, seed=n)
    got = mp_isdf.grid_coulomb_metric(theta, xyz, w)
    want = oracle.grid_coulomb_metric(theta, xyz, w)
    np.testing.assert_array_equal(got, got.T)
    close(got, want, 1e-11)


def test_cross_metric_matches_oracle(mp_isdf):
    # 20011 points (not a multiple of 4: scalar loads in the product), 130
    # right-hand vectors (two column tiles)
    xyz, w, left, right = random_case(20011, 5, 130, seed=4)
    close(mp_isdf.grid_cross_metric(left, right, xyz, w), oracle.grid_cross_metric(left, right, xyz, w), 1e-11)


@pytest.mark.parametrize("block_rows,ld", [(1024, 3001), (128, 3004), (4096, 3004), (0, 3008)])
def test_device_metric_row_blocks_match_oracle(mp_isdf, block_rows, ld):
    """The device entry point with explicit row blocks (3001 points: three
    1024-row blocks, 24 blocks of 128, one partial 4096-row block, automatic)
    and padded operand rows."""
    import torch
    from mp_isdf_amd import capi
    n, kl, kr = 3001, 7, 130
    xyz, w, left, right = random_case(n, kl, kr, seed=9)
    lpad = np.zeros((kl, ld))
    rpad = np.zeros((kr, ld))
    lpad[:, :n] = left
    rpad[:, :n] = right
    dl, dr = torch.tensor(lpad, device="cuda"), torch.tensor(rpad, device="cuda")
    dx, dw = torch.tensor(xyz, device="cuda"), torch.tensor(w, device="cuda")
    out = torch.full((kl, kr), float("nan"), dtype=torch.float64, device="cuda")
    capi.call("mpisdf_grid_coulomb_metric_device", dl, kl, ld, dr, kr, ld, dx, dw, n, False, out, block_rows, None)
    close(out.cpu().numpy(), oracle.grid_cross_metric(left, right, xyz, w), 1e-11)
    with pytest.raises(ValueError, match="block_rows"):
        capi.call("mpisdf_grid_coulomb_metric_device", dl, kl, ld, dr, kr, ld, dx, dw, n, False, out, 100, None)


@pytest.mark.parametrize("block_rows", [128, 1024, 0])
def test_device_symmetric_metric_block_pairs_match_oracle(mp_isdf, block_rows):
    """One vector set, symmetrised: only the kernel blocks on and above the
    diagonal are generated (3001 points: 24 x 24 blocks of 128, 3 x 3 of 1024,
    one automatic block)."""
    import torch
    from mp_isdf_amd import capi
    n, k, ld = 3001, 130, 3004
    xyz, w, theta, _ = random_case(n, k, 1, seed=12)
    tpad = np.zeros((k, ld))
    tpad[:, :n] = theta
    dt = torch.tensor(tpad, device="cuda")
    dx, dw = torch.tensor(xyz, device="cuda"), torch.tensor(w, device="cuda")
    out = torch.full((k, k), float("nan"), dtype=torch.float64, device="cuda")
    capi.call("mpisdf_grid_coulomb_metric_device", dt, k, ld, dt, k, ld, dx, dw, n, True, out, block_rows, None)
    got = out.cpu().numpy()
    np.testing.assert_array_equal(got, got.T)
    close(got, oracle.grid_coulomb_metric(theta, xyz, w), 1e-11)


def test_metric_rejects_zero_weights_and_bad_shapes(mp_isdf):
    xyz, w, theta, _ = random_case(50, 3, 1, seed=1)
    w[4] = 0.0
    with pytest.raises(ValueError, match="nonzero"):
        mp_isdf.grid_coulomb_metric(theta, xyz, w)
    with pytest.raises(ValueError):
        mp_isdf.grid_coulomb_metric(theta[:, :10], xyz, np.abs(w) + 1)
    empty = mp_isdf.grid_coulomb_metric(np.zeros((3, 0)), np.zeros((0, 3)), np.zeros(0))
    np.testing.assert_array_equal(empty, np.zeros((3, 3)))


@pytest.mark.parametrize("k,nao", [(50, 37), (3, 2), (260, 140)])
def test_jk_matches_oracle(mp_isdf, k, nao):
    rng = np.random.default_rng(k)
    x = rng.normal(size=(k, nao))
    z = rng.normal(size=(k, k))
    z = z + z.T
    dm = rng.normal(size=(nao, nao))
    dm = dm + dm.T
    j, kk = mp_isdf.isdf_jk(x, z, dm)
    jo, ko = oracle.evaluate_isdf_jk(x, z, dm)
    close(j, jo, 1e-12)
    close(kk, ko, 1e-12)
    np.testing.assert_array_equal(j, j.T)
    np.testing.assert_array_equal(kk, kk.T)


def test_c1_jk_against_oracle_and_exact_integrals(mp_isdf):
    mol = hydrogen_molecule("sto-3g")
    result = driver.build_isdf(mol, level=3, anchor_rank=3, anchor_algorithm="qrcp_dense", compute_metric=True,
                               metric_algorithm="grid")
    assert len(result["anchor_indices"]) == 3
    assert result["timings"]["coulomb_metric"] >= 0
    theta, coords, weights = result["interp_vecs"], result["coordinates"], result["weights"]
    aw = result["anchor_weights"]
    want_metric = aw[:, None] * oracle.grid_coulomb_metric(theta, coords, weights) * aw[None, :]
    close(result["coulomb_metric"], want_metric, 1e-11)

    funcs = oracle.s_functions(mol.atm, mol.bas, mol.env)
    s12 = oracle.s_overlap(funcs)[0, 1]
    c = np.array([1.0, 1.0]) / np.sqrt(2 * (1 + s12))
    dm = 2 * np.outer(c, c)
    j, k = driver.evaluate_isdf_jk(result, dm)
    jo, ko = oracle.evaluate_isdf_jk(result["X1"], result["coulomb_metric"], dm)
    close(j, jo, 1e-12)
    close(k, ko, 1e-12)
    eri = oracle.s_eri(funcs)
    dj = np.abs(j - np.einsum("uvwx,wx->uv", eri, dm)).max()
    dk = np.abs(k - np.einsum("uwvx,wx->uv", eri, dm)).max()
    print(f"C1 H2/STO-3G level 3: max|dJ| = {dj:.3e}, max|dK| = {dk:.3e}")
    assert dj < 2e-3 and dk < 2e-3  # 1.0e-3 measured on the level-3 Lebedev grid

    # Validation messages of python/isdf.py:140-147
    with pytest.raises(ValueError, match="density"):
        driver.evaluate_isdf_jk(result, np.ones((3, 3)))
    with pytest.raises(ValueError, match="symmetric"):
        driver.evaluate_isdf_jk(result, np.array([[1., 1.], [0., 1.]]))
    with pytest.raises(ValueError, match="metric"):
        driver.evaluate_isdf_jk({**result, "coulomb_metric": None}, dm)
    with pytest.raises(ValueError, match="symmetric"):
        mp_isdf.isdf_jk(result["X1"], result["coulomb_metric"], np.array([[1., 1.], [0., 1.]]))


def test_water_driver_metric_and_jk_match_oracle(mp_isdf):
    """Water cc-pVDZ (s, p and d shells), level-2 grid (20,272 points), 60
    qrcp_dense anchors: the driver's metric over the real interpolation
    vectors and J/K of a symmetric density equal the oracle's."""
    mol = molecule.single_water("cc-pvdz")
    result = driver.build_isdf(mol, level=2, anchor_rank=60, anchor_algorithm="qrcp_dense", compute_metric=True,
                               metric_algorithm="grid")
    theta, coords, weights = result["interp_vecs"], result["coordinates"], result["weights"]
    aw = result["anchor_weights"]
    want = aw[:, None] * oracle.grid_coulomb_metric(theta, coords, weights) * aw[None, :]
    close(result["coulomb_metric"], want, 1e-11)
    rng = np.random.default_rng(3)
    c = rng.normal(size=(mol.nao, 5))
    dm = 2 * c @ c.T
    j, k = driver.evaluate_isdf_jk(result, dm)
    jo, ko = oracle.evaluate_isdf_jk(result["X1"], result["coulomb_metric"], dm)
    close(j, jo, 1e-12)
    close(k, ko, 1e-12)
