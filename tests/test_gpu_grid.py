"""The Becke partition on the GPU evaluates the same function as the host loop
with no FMA contraction, so the device weights equal the host weights bit for
bit (the qrcp_online fixture's anchors depend on the exact grid)."""
import numpy as np
import pytest

from mp_isdf_amd import grid as gridmod
from mp_isdf_amd import molecule

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nwater, basis, level", [(8, "cc-pvdz", 3), (16, "6-31g*", 2), (1, "sto-3g", 4)])
def test_device_partition_equals_host(mp_isdf, nwater, basis, level):
    mol = molecule.water_cluster(nwater, basis) if nwater > 1 else molecule.single_water(basis)
    xd, wd = gridmod.make_grid(mol, level, device=True)
    xh, wh = gridmod.make_grid(mol, level, device=False)
    np.testing.assert_array_equal(xd, xh)
    np.testing.assert_array_equal(wd, wh)
    assert np.all(np.isfinite(wd)) and np.count_nonzero(wd) > 0.9 * len(wd)
    # The 74- and 266-point Lebedev rules carry negative weights (levels 0 and
    # 2-3 pruned shells), so some grid weights are negative, as in PySCF's
    # grids; the weights still integrate a positive density.
    assert np.sum(wd) > 0 and np.sum(wd[wd < 0]) > -0.1 * np.sum(wd[wd > 0])


def test_higher_levels_raise(mp_isdf):
    with pytest.raises(ValueError, match="770-point Lebedev rule"):
        gridmod.make_grid(molecule.single_water("sto-3g"), 5)
