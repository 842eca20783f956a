"""Molecular grids for the selection workloads (PySCF-free front end).

``make_grid(mol, level)`` returns (coords[n,3] Bohr, weights[n]) from the
library's Becke/Treutler generator; when PySCF is importable the caller may
pass its ``Grids`` arrays instead (the reference takes them from PySCF,
python/isdf.py:488-496).
"""
from __future__ import annotations

import numpy as np

from .basis import CHARGES, Molecule


def make_grid(mol: Molecule, level: int = 3, device: bool = True):
    """Becke-partitioned grid; ``device`` runs the partition on the GPU when one
    is present (the same weights as the host loop)."""
    from . import mp_isdf  # compiled extension
    charges = np.asarray([CHARGES[s] for s in mol.symbols], dtype=np.int32)
    return mp_isdf.becke_grid(charges, np.ascontiguousarray(mol.coords), level, device)
