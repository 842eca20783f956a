"""GPU parity of AO evaluation against the FP64 oracle (reference semantics of
cpp/src/ao_eval.cpp:164-252; tolerances of python/test/test_ao_eval.py:33,49:
rtol 1e-11, atol 1e-12 for FP64)."""
import ctypes

import numpy as np
import pytest

import oracle
from mp_isdf_amd import basis, molecule

pytestmark = pytest.mark.gpu


def check(mp_isdf, mol, pts, cart=False):
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=cart)
    assert ctx.nao == mol.nao
    np.testing.assert_allclose(ctx.eval_ao(pts), oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, cart=cart),
                               rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(ctx.eval_ao_gradient(pts),
                               oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, deriv=1, cart=cart),
                               rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("basis_name", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_water_values_and_gradients(mp_isdf, basis_name):
    pts = np.vstack(([0.0, 0.0, 0.0], np.random.default_rng(42).normal(size=(17, 3))))
    check(mp_isdf, molecule.single_water(basis_name), pts)


def test_cartesian_d(mp_isdf):
    pts = np.random.default_rng(3).normal(size=(50, 3)) * 2
    check(mp_isdf, molecule.single_water("6-31g*", cart=True), pts, cart=True)


def test_three_s_shell_table(mp_isdf):
    atm, bas, env = basis.three_s_shell_tables()
    pts = np.random.default_rng(23).normal(size=(17, 3))
    got = mp_isdf.eval_ao(atm, bas, env, pts)
    centres = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]], float)
    expect = np.stack([np.exp(-a * np.sum((pts - c) ** 2, 1)) / np.sqrt(4 * np.pi)
                       for a, c in zip((0.8, 1.1, 1.3), centres)], 1)
    np.testing.assert_allclose(got, expect, rtol=1e-13, atol=1e-16)


def test_cluster_many_tiles_with_culling(mp_isdf):
    mol = molecule.water_cluster(8)
    rng = np.random.default_rng(9)
    # Points spread over and beyond the cluster exercise per-tile culling.
    pts = rng.uniform(-14, 14, size=(3000, 3))
    ctx = mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env)
    ref = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    got = ctx.eval_ao(pts)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(ctx.eval_ao_gradient(pts[:700]),
                               oracle.ao_eval(mol.atm, mol.bas, mol.env, pts[:700], deriv=1),
                               rtol=1e-11, atol=1e-12)


def test_context_owns_tables_and_empty_grid(mp_isdf):
    mol = molecule.single_water("cc-pvdz")
    atm, bas, env = mol.atm.copy(), mol.bas.copy(), mol.env.copy()
    pts = np.random.default_rng(7).normal(size=(9, 3))
    expected = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    ctx = mp_isdf.LibcintContext(atm, bas, env)
    atm.fill(0); bas.fill(0); env.fill(0)
    for _ in range(2):
        np.testing.assert_allclose(ctx.eval_ao(pts), expected, rtol=1e-11, atol=1e-12)
    assert ctx.eval_ao(np.empty((0, 3))).shape == (0, 24)
    assert ctx.eval_ao_gradient(np.empty((0, 3))).shape == (3, 0, 24)
    assert mp_isdf.eval_ao(mol.atm, mol.bas, mol.env, np.empty((0, 3))).shape == (0, 24)
    for space in ("auto", "device"):
        np.testing.assert_allclose(ctx.eval_ao(pts, space=space), expected, rtol=1e-11, atol=1e-12)


def test_reduced_precision_device_outputs(mp_isdf):
    """FP32 / BF16 outputs of the device entry point against the FP64 oracle."""
    import torch
    import mp_isdf_amd
    mol = molecule.water_cluster(8)
    lib = ctypes.CDLL(str(mp_isdf_amd.LIBRARY))
    ctx = ctypes.c_void_p()
    assert lib.mpisdf_context_create(mol.atm.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(mol.atm)),
                                     mol.bas.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(mol.bas)),
                                     mol.env.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(len(mol.env)),
                                     0, ctypes.byref(ctx)) == 0
    pts = np.random.default_rng(1).uniform(-10, 10, size=(2048, 3))
    ref = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    xyz = torch.tensor(pts, device="cuda")
    for dtype, code, tol in ((torch.float32, 1, 2e-6), (torch.bfloat16, 2, 8e-3)):
        out = torch.empty((len(pts), mol.nao), dtype=dtype, device="cuda")
        rc = lib.mpisdf_ao_eval_device(ctx, ctypes.c_void_p(xyz.data_ptr()), ctypes.c_int64(len(pts)), 0, code,
                                       ctypes.c_void_p(out.data_ptr()), ctypes.c_int64(mol.nao),
                                       ctypes.c_int64(len(pts) * mol.nao), None)
        assert rc == 0
        torch.cuda.synchronize()
        got = out.double().cpu().numpy()
        scale = np.abs(ref).max()
        assert np.max(np.abs(got - ref)) <= tol * scale + 1e-9
    lib.mpisdf_context_destroy(ctx)
