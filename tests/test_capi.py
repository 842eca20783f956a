"""CPU checks of the boundary: the C-ABI library loads and exports every
symbol include/mp_isdf_amd.h declares; the pybind11 module has the reference
surface; host-space requests and missing devices fail loudly."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import mp_isdf_amd

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "mp_isdf_amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mpisdf_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(mp_isdf_amd.LIBRARY))
    names = declared_symbols()
    assert "mpisdf_select" in names and "mpisdf_eval_ao" in names
    for name in names:
        assert hasattr(lib, name), name


def test_module_surface_matches_reference():
    m = mp_isdf_amd.mp_isdf
    for name in ("LibcintContext", "AnchorSelection", "DistributedAnchorSelection", "eval_ao",
                 "eval_ao_gradient", "cuda_enabled", "device_enabled", "rbrp_ao_device",
                 # the driver's GPU counterparts (python/isdf.py:48-151) and the rbrp_ao_cpu surface
                 "grid_coulomb_metric", "grid_cross_metric", "isdf_jk", "rbrp_ao", "isdf_fit"):
        assert hasattr(m, name), name
    assert hasattr(m.LibcintContext, "overlap")
    assert [v for v in m.AxisType.__members__] == [
        "AtomicOrbital", "AtomicOrbitalGradX", "AtomicOrbitalGradY", "AtomicOrbitalGradZ", "OccupiedOrbital"]
    assert list(m.LocalAnchorSelectionAlgorithm.__members__) == [
        "QRColumnPivoting", "OnlineQRColumnPivoting", "RandomizedQuadratureKhatriRao", "CentroidalVoronoi",
        "RandomizedAODensityKhatriRao", "RandomizedAODensityNormalizedKhatriRao"]
    assert list(m.DistributedAnchorSelectionAlgorithm.__members__) == ["SVD", "RandomizedQRColumnPivoting"]


def test_host_space_and_bad_space_raise_value_error():
    from mp_isdf_amd.molecule import single_water
    m = mp_isdf_amd.mp_isdf
    mol = single_water("sto-3g")
    with pytest.raises(ValueError):
        m.eval_ao(mol.atm, mol.bas, mol.env, np.zeros((2, 3)), space="host")
    with pytest.raises(ValueError):
        m.eval_ao(mol.atm, mol.bas, mol.env, np.zeros((2, 3)), space="invalid")


def test_table_shape_validation():
    m = mp_isdf_amd.mp_isdf
    with pytest.raises(ValueError):
        m.LibcintContext(np.zeros((1, 5), np.int32), np.zeros((1, 8), np.int32), np.zeros(30))


@pytest.mark.skipif(mp_isdf_amd.mp_isdf.device_enabled, reason="checks the no-device path")
def test_no_device_fails_loudly():
    from mp_isdf_amd.molecule import single_water
    mol = single_water("sto-3g")
    with pytest.raises(RuntimeError, match="no HIP device"):
        mp_isdf_amd.mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
