"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from
closed forms, scipy and NumPy) checked against the oracle on the CPU and
against the HIP path on the GPU."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from mp_isdf_amd import basis

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def test_golden_three_s_ao_oracle():
    z = load("ao_three_s.npz")
    atm, bas, env = basis.three_s_shell_tables()
    np.testing.assert_allclose(oracle.ao_eval(atm, bas, env, z["points"]), z["values"], rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(oracle.ao_eval(atm, bas, env, z["points"], deriv=1), z["gradient"], rtol=1e-13,
                               atol=1e-16)


def test_golden_qrcp_oracle():
    z = load("qrcp_three_s.npz")
    _, piv, _ = oracle.geqp3(z["pair_matrix"])
    np.testing.assert_array_equal(piv[:len(z["scipy_pivots"])], z["scipy_pivots"])
    atm, bas, env = basis.three_s_shell_tables()
    got = oracle.select(atm, bas, env, z["points"], z["weights"], "ao", max_rank=4)
    np.testing.assert_array_equal(got, z["select_rank4"])
    got = oracle.select(atm, bas, env, z["points"], z["weights"], "ao", tolerance=1e-6)
    np.testing.assert_array_equal(got, z["select_tol_1e6"])


def test_golden_khatri_rao_oracle():
    z = load("khatri_rao.npz")
    y = oracle.sketch(z["phi"], z["u"], z["v"], z["sqrt_weights"])
    np.testing.assert_allclose(y, z["y"], rtol=1e-13, atol=1e-15)


def test_golden_pair_error_kat():
    kat = json.loads((GOLDEN / "pair_error_kat.json").read_text())
    err = oracle.pair_space_error_norms(np.array(kat["exact_pairs"]), np.array(kat["anchor_pairs"]),
                                        np.array(kat["anchor_weights"]), np.array(kat["theta"]),
                                        np.array(kat["grid_weights"]), block_size=1)
    np.testing.assert_allclose(err["l2"], kat["l2"], rtol=1e-15)
    np.testing.assert_allclose(err["relative"], kat["relative"], rtol=1e-15)


@pytest.mark.gpu
def test_golden_three_s_ao_gpu(mp_isdf):
    z = load("ao_three_s.npz")
    atm, bas, env = basis.three_s_shell_tables()
    ctx = mp_isdf.LibcintContext(atm, bas, env)
    np.testing.assert_allclose(ctx.eval_ao(z["points"]), z["values"], rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(ctx.eval_ao_gradient(z["points"]), z["gradient"], rtol=1e-12, atol=1e-16)


@pytest.mark.gpu
def test_golden_qrcp_gpu(mp_isdf):
    z = load("qrcp_three_s.npz")
    atm, bas, env = basis.three_s_shell_tables()
    sel = mp_isdf.AnchorSelection(mp_isdf.LibcintContext(atm, bas, env))
    ao = mp_isdf.AxisType.AtomicOrbital
    for alg in (mp_isdf.LocalAnchorSelectionAlgorithm.QRColumnPivoting,
                mp_isdf.LocalAnchorSelectionAlgorithm.OnlineQRColumnPivoting):
        assert sel.select(z["points"], z["weights"], ao, max_rank=4, algorithm=alg) == list(z["select_rank4"])
        assert sel.select(z["points"], z["weights"], ao, tolerance=1e-6, algorithm=alg) == list(z["select_tol_1e6"])


@pytest.mark.gpu
def test_golden_khatri_rao_gpu(mp_isdf):
    import torch
    from mp_isdf_amd import capi
    z = load("khatri_rao.npz")
    n, d = z["phi"].shape
    ell = z["u"].shape[1]
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    y = torch.zeros((n, ell), dtype=torch.float64, device="cuda")
    capi.call("mpisdf_kr_sketch_f64_device", dev(z["phi"]), n, d, dev(z["u"]), dev(z["v"]), ell,
              dev(z["sqrt_weights"]), y, ell, None, None)
    np.testing.assert_allclose(y.cpu().numpy(), z["y"], rtol=1e-12, atol=1e-14)
