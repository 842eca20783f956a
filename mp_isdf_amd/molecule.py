"""Seeded synthetic systems for the selection workloads.

The reference ships no water clusters (python/xyz/), so the water systems are
built here by seeded rejection sampling:

* rigid water: O-H 0.957 Angstrom, angle H-O-H 104.5 degrees;
* orientations: uniform random rotations (unit quaternions, Shoemake 1992);
* oxygen positions: uniform in a cube centred at the origin, rejecting any
  O-O distance below 2.7 Angstrom;
* random stream: ``numpy.random.default_rng(nwater)`` unless a seed is given.

Box edges follow the workload definitions: 9 A (8 waters), 11.5 A (16),
14 A (32), 18 A (64).  Coordinates are returned in Bohr.
"""
from __future__ import annotations

import numpy as np

from .basis import BOHR, Molecule, make_tables

WATER_BOX_ANGSTROM = {8: 9.0, 16: 11.5, 32: 14.0, 64: 18.0}
OH_ANGSTROM = 0.957
HOH_DEGREES = 104.5
MIN_OO_ANGSTROM = 2.7


def _random_rotation(rng) -> np.ndarray:
    u1, u2, u3 = rng.random(3)
    a = np.sqrt(1.0 - u1)
    b = np.sqrt(u1)
    w, x, y, z = (a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3))
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster_angstrom(nwater: int, box: float | None = None, seed: int | None = None,
                           max_tries: int = 1_000_000):
    """Return (symbols, coords[3*nwater, 3] in Angstrom) for a seeded cluster."""
    if nwater < 1:
        raise ValueError("nwater must be positive")
    if box is None:
        box = WATER_BOX_ANGSTROM.get(nwater, 18.0 * (nwater / 64.0) ** (1.0 / 3.0))
    rng = np.random.default_rng(nwater if seed is None else seed)
    half = np.radians(HOH_DEGREES) / 2
    monomer = np.array([
        [0.0, 0.0, 0.0],
        [OH_ANGSTROM * np.sin(half), 0.0, OH_ANGSTROM * np.cos(half)],
        [-OH_ANGSTROM * np.sin(half), 0.0, OH_ANGSTROM * np.cos(half)],
    ])
    oxygens = []
    tries = 0
    while len(oxygens) < nwater:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place waters; enlarge the box")
        candidate = (rng.random(3) - 0.5) * box
        if oxygens and np.min(np.linalg.norm(np.asarray(oxygens) - candidate, axis=1)) < MIN_OO_ANGSTROM:
            continue
        oxygens.append(candidate)
    symbols, coords = [], []
    for o in oxygens:
        rot = _random_rotation(rng)
        for sym, r in zip(("O", "H", "H"), monomer):
            symbols.append(sym)
            coords.append(o + rot @ r)
    return symbols, np.asarray(coords)


def water_cluster(nwater: int, basis: str = "cc-pvdz", cart: bool = False,
                  box: float | None = None, seed: int | None = None) -> Molecule:
    symbols, xyz = water_cluster_angstrom(nwater, box=box, seed=seed)
    return make_tables(symbols, xyz / BOHR, basis=basis, cart=cart)


def hydrogen_molecule(basis: str = "sto-3g", half_bond_angstrom: float = 0.37) -> Molecule:
    """H2 with atoms at +-half_bond on z (0.37 A gives the 0.74 A bond)."""
    xyz = np.array([[0.0, 0.0, -half_bond_angstrom], [0.0, 0.0, half_bond_angstrom]])
    return make_tables(["H", "H"], xyz / BOHR, basis=basis)


def single_water(basis: str = "cc-pvdz", cart: bool = False) -> Molecule:
    """The water geometry used by the reference AO tests (Angstrom input):
    O 0 0 0; H 0 0.75 0.5; H 0 -0.75 0.5 (python/test/test_ao_eval.py:25-27)."""
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.75, 0.5], [0.0, -0.75, 0.5]])
    return make_tables(["O", "H", "H"], xyz / BOHR, basis=basis, cart=cart)
