"""Nuclear gradients of the fitted pair-reconstruction loss and of the
Coulomb energy of the fitted density.

The reference has no nuclear derivatives.  It provides the grid-coordinate AO
derivatives (dx/dy/dz axes, cpp/src/ao_eval.cpp:215-244), which the adjoint
here is checked against.

The loss is the squared l2 pair-space error of the FP64 ISDF fit:

    L(R) = sum_g |w_g| || phi_g phi_g^T - sum_k c_gk phi_k phi_k^T ||_F^2,

where c_g are the least-squares coefficients on the anchors.  Grid points,
weights and anchor points are held fixed; the nuclear coordinates R move the
atom-centred AOs, d phi_mu(r) / d R_A = -grad_r phi_mu(r) for mu on atom A.
Because c is optimal, dL/dR is the partial derivative at fixed c (envelope
theorem), so the backward pass needs no derivative of the solve.

``PairLossFunction`` is a ``torch.autograd.Function``:

* forward: the GPU fit and its loss (no derivative blocks kept);
* backward: the hand-written adjoint on the GPU (``mp_isdf.isdf_fit_gradient``),
  which re-evaluates AO values and gradients tile by tile instead of storing
  them.

The Coulomb energy is the ISDF J energy of a fixed symmetric AO density
matrix D, E_J = tr(J D) / 2 with J from ``evaluate_isdf_jk``
(python/isdf.py:130-151) and the grid Coulomb metric of the refit
(isdf.py:48-67, :674-689).  With rho_p = phi_p^T D phi_p at the anchors and
u = C^T rho the fitted density on the grid, E_J = u^T W K W u / 2.  The
coefficients C = G2^-1 S are not a minimiser of E_J, so its adjoint carries
the derivative of the solve: with lam = G2^-1 rho, v = W K W u and
beta = G2^-1 S v,

    dE_J = sum_g v_g lam^T ds_g + beta^T (d rho - dG2 lam).

``CoulombEnergyFunction`` wraps it as a ``torch.autograd.Function``
(``mp_isdf.coulomb_energy_gradient``).
"""
from __future__ import annotations

import numpy as np
import torch

from .basis import Molecule


def with_coords(mol: Molecule, coords) -> Molecule:
    """The same molecule and basis with nuclei at ``coords`` (Bohr, [natm, 3])."""
    coords = np.ascontiguousarray(coords, dtype=np.float64).reshape(mol.natm, 3)
    env = mol.env.copy()
    for i in range(mol.natm):
        p = int(mol.atm[i, 1])
        env[p:p + 3] = coords[i]
    return Molecule(list(mol.symbols), coords, mol.basis, mol.cart, mol.atm.copy(), mol.bas.copy(), env)


def _context(mol: Molecule):
    from . import mp_isdf
    return mp_isdf.LibcintContext(mol.atm, mol.bas, mol.env, cart=mol.cart)


def pair_loss_value(mol: Molecule, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size: int = 4096) -> float:
    """L at the molecule's current geometry (forward only)."""
    from . import mp_isdf
    fit = mp_isdf.isdf_fit(_context(mol), anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size=batch_size)
    return float(fit["pair_space_error"]["l2"]) ** 2


def pair_loss_and_gradient(mol: Molecule, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size: int = 4096):
    """(L, dL/dR [natm, 3], per-AO adjoint [3, nao]) at the molecule's geometry."""
    from . import mp_isdf
    out = mp_isdf.isdf_fit_gradient(_context(mol), np.ascontiguousarray(anchor_xyz, dtype=np.float64),
                                    np.ascontiguousarray(anchor_w, dtype=np.float64),
                                    np.ascontiguousarray(grid_xyz, dtype=np.float64),
                                    np.ascontiguousarray(grid_w, dtype=np.float64), batch_size=batch_size)
    return out["loss"], out["grad_atoms"], out["grad_ao"]


class PairLossFunction(torch.autograd.Function):
    """L(R) with a hand-written GPU adjoint; R is a [natm, 3] float64 tensor."""

    @staticmethod
    def forward(ctx, coords, mol, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size=4096):
        geometry = with_coords(mol, coords.detach().cpu().numpy())
        ctx.saved = (geometry, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size, coords.device)
        value = pair_loss_value(geometry, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size)
        return torch.tensor(value, dtype=torch.float64, device=coords.device)

    @staticmethod
    def backward(ctx, grad_output):
        geometry, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size, device = ctx.saved
        _, grad, _ = pair_loss_and_gradient(geometry, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size)
        g = torch.as_tensor(grad, dtype=torch.float64, device=device)
        return grad_output * g, None, None, None, None, None, None


def pair_loss(coords: torch.Tensor, mol: Molecule, anchor_xyz, anchor_w, grid_xyz, grid_w,
              batch_size: int = 4096) -> torch.Tensor:
    """Differentiable pair-reconstruction loss of the fit on fixed anchors and grid."""
    return PairLossFunction.apply(coords, mol, anchor_xyz, anchor_w, grid_xyz, grid_w, batch_size)


def coulomb_energy_value(mol: Molecule, anchor_xyz, grid_xyz, grid_w, dm, batch_size: int = 4096) -> float:
    """E_J at the molecule's current geometry (forward only)."""
    from . import mp_isdf
    out = mp_isdf.coulomb_energy_gradient(_context(mol), np.ascontiguousarray(anchor_xyz, dtype=np.float64),
                                          np.ascontiguousarray(grid_xyz, dtype=np.float64),
                                          np.ascontiguousarray(grid_w, dtype=np.float64),
                                          np.ascontiguousarray(dm, dtype=np.float64), batch_size=batch_size,
                                          gradient=False)
    return out["energy"]


def coulomb_energy_and_gradient(mol: Molecule, anchor_xyz, grid_xyz, grid_w, dm, batch_size: int = 4096):
    """(E_J, dE_J/dR [natm, 3], per-AO adjoint [3, nao]) at the molecule's geometry."""
    from . import mp_isdf
    out = mp_isdf.coulomb_energy_gradient(_context(mol), np.ascontiguousarray(anchor_xyz, dtype=np.float64),
                                          np.ascontiguousarray(grid_xyz, dtype=np.float64),
                                          np.ascontiguousarray(grid_w, dtype=np.float64),
                                          np.ascontiguousarray(dm, dtype=np.float64), batch_size=batch_size)
    return out["energy"], out["grad_atoms"], out["grad_ao"]


class CoulombEnergyFunction(torch.autograd.Function):
    """E_J(R) with a hand-written GPU adjoint; R is a [natm, 3] float64 tensor."""

    @staticmethod
    def forward(ctx, coords, mol, anchor_xyz, grid_xyz, grid_w, dm, batch_size=4096):
        geometry = with_coords(mol, coords.detach().cpu().numpy())
        ctx.saved = (geometry, anchor_xyz, grid_xyz, grid_w, dm, batch_size, coords.device)
        value = coulomb_energy_value(geometry, anchor_xyz, grid_xyz, grid_w, dm, batch_size)
        return torch.tensor(value, dtype=torch.float64, device=coords.device)

    @staticmethod
    def backward(ctx, grad_output):
        geometry, anchor_xyz, grid_xyz, grid_w, dm, batch_size, device = ctx.saved
        _, grad, _ = coulomb_energy_and_gradient(geometry, anchor_xyz, grid_xyz, grid_w, dm, batch_size)
        g = torch.as_tensor(grad, dtype=torch.float64, device=device)
        return grad_output * g, None, None, None, None, None, None


def coulomb_energy(coords: torch.Tensor, mol: Molecule, anchor_xyz, grid_xyz, grid_w, dm,
                   batch_size: int = 4096) -> torch.Tensor:
    """Differentiable Coulomb energy of the ISDF fit on fixed anchors and grid for the density matrix dm."""
    return CoulombEnergyFunction.apply(coords, mol, anchor_xyz, grid_xyz, grid_w, dm, batch_size)
