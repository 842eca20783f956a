"""Pins the oracle's AO evaluator (oracle/ao_oracle.cpp) without PySCF.

Checks are chosen so a dropped term, wrong sign/index or transposed operand
fails one of them: closed forms, sphere orthonormality of the harmonics,
radial normalisation integrals, finite differences, and published constants
of real solid harmonics (libcint's d/f table entries are these closed forms).
"""
import numpy as np
import pytest

import oracle
from mp_isdf_amd import basis, molecule


def sphere_grid(nt=24, nph=48):
    x, wx = np.polynomial.legendre.leggauss(nt)
    th = np.arccos(x)
    ph = 2 * np.pi * np.arange(nph) / nph
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    W = np.outer(wx, np.full(nph, 2 * np.pi / nph))
    dirs = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    return dirs.reshape(-1, 3), W.reshape(-1)


def monomials(dirs, l):
    cols = []
    for lx in range(l, -1, -1):
        for ly in range(l - lx, -1, -1):
            lz = l - lx - ly
            cols.append(dirs[:, 0] ** lx * dirs[:, 1] ** ly * dirs[:, 2] ** lz)
    return np.stack(cols, -1)


def test_three_s_shells_closed_form():
    atm, bas, env = basis.three_s_shell_tables()
    pts = np.random.default_rng(23).normal(size=(17, 3))
    ao = oracle.ao_eval(atm, bas, env, pts)
    centres = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]], float)
    expect = np.stack([np.exp(-a * np.sum((pts - c) ** 2, 1)) / np.sqrt(4 * np.pi)
                       for a, c in zip((0.8, 1.1, 1.3), centres)], 1)
    np.testing.assert_allclose(ao, expect, rtol=1e-14, atol=1e-16)
    grad = oracle.ao_eval(atm, bas, env, pts, deriv=1)
    gexp = np.stack([-2 * a * (pts - c) * e[:, None] for a, c, e in
                     zip((0.8, 1.1, 1.3), centres, expect.T)], 1)  # (n, 3 shells, 3 xyz)
    np.testing.assert_allclose(grad, np.transpose(gexp, (2, 0, 1)), rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("l", range(0, 6))
def test_harmonics_orthonormal_on_sphere(l):
    dirs, w = sphere_grid()
    y = monomials(dirs, l) @ oracle.c2s(l).T
    gram = (y * w[:, None]).T @ y
    np.testing.assert_allclose(gram, np.eye(2 * l + 1), atol=1e-13)


def test_harmonic_constants_and_order():
    # p: x, y, z with sqrt(3/4pi); d: m=-2..2 = xy, yz, z^2, xz, x^2-y^2
    s3 = np.sqrt(3 / (4 * np.pi))
    np.testing.assert_allclose(oracle.c2s(1), s3 * np.eye(3), rtol=1e-15)
    d = oracle.c2s(2)  # cart order xx xy xz yy yz zz
    c15 = np.sqrt(15 / (4 * np.pi))
    c5 = np.sqrt(5 / (16 * np.pi))
    expect = np.zeros((5, 6))
    expect[0, 1] = c15
    expect[1, 4] = c15
    expect[2, [0, 3, 5]] = [-c5, -c5, 2 * c5]
    expect[3, 2] = c15
    expect[4, [0, 3]] = [c15 / 2, -c15 / 2]
    np.testing.assert_allclose(d, expect, rtol=1e-14, atol=1e-15)
    f = oracle.c2s(3)  # f-3 = sqrt(35/32pi) y(3x^2 - y^2): xxy, yyy entries
    c = np.sqrt(35 / (32 * np.pi))
    np.testing.assert_allclose(f[0, [1, 6]], [3 * c, -c], rtol=1e-14)
    np.testing.assert_allclose(f[6, [0, 3]], [c, -3 * c], rtol=1e-14)  # f+3 = x(x^2 - 3y^2)


@pytest.mark.parametrize("basis_name", ["sto-3g", "cc-pvdz", "6-31g*"])
def test_single_atom_overlap_is_identity(basis_name):
    mol = basis.make_tables(["O"], np.zeros((1, 3)), basis_name)
    dirs, w = sphere_grid(30, 60)
    r, wr = np.polynomial.legendre.leggauss(300)
    r = (r + 1) / 2 * 14.0
    wr = wr * 7.0
    pts = (r[:, None, None] * dirs[None]).reshape(-1, 3)
    wts = (wr[:, None] * r[:, None] ** 2 * w[None]).reshape(-1)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    s = ao.T @ (ao * wts[:, None])
    # Same-shell overlaps are exact; the second O s contraction overlaps the first.
    np.testing.assert_allclose(np.diag(s), 1.0, atol=2e-10)


def test_cartesian_d_norms():
    mol = basis.make_tables(["O"], np.zeros((1, 3)), "6-31g*", cart=True)
    assert oracle.nao(mol.atm, mol.bas, mol.env, cart=True) == 15
    dirs, w = sphere_grid(30, 60)
    r, wr = np.polynomial.legendre.leggauss(300)
    r = (r + 1) / 2 * 14.0
    wr = wr * 7.0
    pts = (r[:, None, None] * dirs[None]).reshape(-1, 3)
    wts = (wr[:, None] * r[:, None] ** 2 * w[None]).reshape(-1)
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, cart=True)
    norms = np.einsum("gi,g,gi->i", ao, wts, ao)
    # libcint cartesian d (no extra factor): x^2 -> 4pi/5, xy -> 4pi/15
    np.testing.assert_allclose(norms[9:15], [4 * np.pi / 5, 4 * np.pi / 15, 4 * np.pi / 15,
                                             4 * np.pi / 5, 4 * np.pi / 15, 4 * np.pi / 5], rtol=1e-9)


@pytest.mark.parametrize("basis_name,cart", [("cc-pvdz", False), ("6-31g*", True), ("sto-3g", False)])
def test_gradient_matches_finite_differences(basis_name, cart):
    mol = molecule.single_water(basis_name, cart=cart)
    pts = np.vstack(([0.0, 0.0, 0.0], np.random.default_rng(11).normal(size=(9, 3))))
    grad = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, deriv=1, cart=cart)
    h = 1e-5
    for axis in range(3):
        dp = pts.copy(); dp[:, axis] += h
        dm = pts.copy(); dm[:, axis] -= h
        fd = (oracle.ao_eval(mol.atm, mol.bas, mol.env, dp, cart=cart) -
              oracle.ao_eval(mol.atm, mol.bas, mol.env, dm, cart=cart)) / (2 * h)
        np.testing.assert_allclose(grad[axis], fd, rtol=1e-6, atol=1e-7)
    both = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, deriv=2, cart=cart)
    np.testing.assert_array_equal(both[1:], grad)
    np.testing.assert_array_equal(both[0], oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, cart=cart))


def test_axis_assembly_and_scales():
    mol = molecule.hydrogen_molecule("sto-3g")
    pts = np.random.default_rng(12).normal(size=(8, 3))
    w = np.array([1., 2., 0., -1., 1., 3., 1., 1.])
    ao = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts)
    grad = oracle.ao_eval(mol.atm, mol.bas, mol.env, pts, deriv=1)
    c = np.array([[0.5, 1.2], [0.5, -1.2]])
    feat = oracle.axis_eval(mol.atm, mol.bas, mol.env, pts, ["dz", "occ", "ao", "dz"], mo_coeff=c, nocc=1)
    np.testing.assert_allclose(feat, np.hstack((grad[2], ao @ c[:, :1], ao, grad[2])), rtol=1e-14)
