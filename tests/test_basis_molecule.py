import numpy as np
import pytest

from mp_isdf_amd import basis, molecule


def test_nao_counts():
    assert molecule.single_water("cc-pvdz").nao == 24
    assert molecule.single_water("sto-3g").nao == 7
    assert molecule.single_water("6-31g*", cart=True).nao == 19
    assert molecule.hydrogen_molecule("sto-3g").nao == 2
    assert molecule.water_cluster(8).nao == 192


def test_contracted_radial_normalisation():
    for l, exps, cols in basis.BASIS_SETS["cc-pvdz"]["O"]:
        c = basis.normalized_coefficients(l, exps, cols)
        e = np.asarray(exps)
        s = basis.gaussian_int(2 * l + 2, e[:, None] + e[None, :])
        np.testing.assert_allclose(np.einsum("pi,pq,qi->i", c, s, c), 1.0, rtol=1e-13)


@pytest.mark.parametrize("nwater", [8, 16])
def test_water_cluster_geometry(nwater):
    sym, xyz = molecule.water_cluster_angstrom(nwater)
    xyz = np.asarray(xyz)
    o = xyz[0::3]
    for i in range(nwater):
        h1, h2 = xyz[3 * i + 1] - o[i], xyz[3 * i + 2] - o[i]
        np.testing.assert_allclose(np.linalg.norm(h1), 0.957, rtol=1e-12)
        np.testing.assert_allclose(np.degrees(np.arccos(h1 @ h2 / 0.957 ** 2)), 104.5, rtol=1e-10)
    d = np.linalg.norm(o[:, None] - o[None], axis=-1) + np.eye(nwater) * 99
    assert d.min() >= 2.7
    sym2, xyz2 = molecule.water_cluster_angstrom(nwater)
    np.testing.assert_array_equal(xyz, xyz2)  # seeded
