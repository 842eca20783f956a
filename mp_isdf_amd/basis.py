"""Gaussian basis tables in libcint ATM/BAS/ENV form, built without PySCF.

The selection path consumes exactly the three arrays PySCF hands to the
reference (``mol._atm``, ``mol._bas``, ``mol._env``; python/isdf.py:501 ->
cpp/src/ao_eval_pybind.cpp:33-46).  This module produces the same layout from
embedded basis-set data so that every configuration runs offline:

* ``atm[natm, 6]`` int32: charge, pointer to the xyz coordinates in ``env``.
* ``bas[nbas, 8]`` int32: atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, 0.
* ``env`` float64: 20 reserved slots, then coordinates, exponents and
  contraction coefficients (``coeff[ctr * nprim + prim]``).

Contraction coefficients follow PySCF's storage convention: each primitive is
scaled by ``1/sqrt(int_0^inf r^(2l+2) exp(-2 a r^2) dr)`` and each contracted
radial function ``r^l sum_p c_p exp(-a_p r^2)`` then has unit norm.  Together
with unit-sphere real spherical harmonics (l=0: 1/sqrt(4 pi), l=1:
sqrt(3/(4 pi)), cpp/src/ao_eval.cpp:135-137) every spherical AO is normalised.

Basis data (exponent, coefficient columns) are the published sets for H and O:
STO-3G, cc-pVDZ (Dunning 1989, general contraction) and 6-31G*.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ATM_SLOTS = 6
BAS_SLOTS = 8
PTR_ENV_START = 20
CHARGE_OF, PTR_COORD = 0, 1
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = 0, 1, 2, 3, 4, 5, 6

BOHR = 0.52917721092  # Angstrom per Bohr (PySCF's param.BOHR)

CHARGES = {"H": 1, "He": 2, "O": 8}

# Each entry: list of (l, exponents, coefficient columns).  A column list with
# more than one entry is a general contraction (nctr > 1) sharing exponents.
_STO3G = {
    "H": [
        (0, [3.42525091, 0.62391373, 0.16885540],
            [[0.15432897, 0.53532814, 0.44463454]]),
    ],
    "He": [
        (0, [6.36242139, 1.15892300, 0.31364979],
            [[0.15432897, 0.53532814, 0.44463454]]),
    ],
    "O": [
        (0, [130.7093200, 23.8088610, 6.4436083],
            [[0.15432897, 0.53532814, 0.44463454]]),
        (0, [5.0331513, 1.1695961, 0.3803890],
            [[-0.09996723, 0.39951283, 0.70011547]]),
        (1, [5.0331513, 1.1695961, 0.3803890],
            [[0.15591627, 0.60768372, 0.39195739]]),
    ],
}

_CCPVDZ = {
    "H": [
        (0, [13.0100000, 1.9620000, 0.4446000, 0.1220000],
            [[0.0196850, 0.1379770, 0.4781480, 0.5012400],
             [0.0, 0.0, 0.0, 1.0]]),
        (1, [0.7270000], [[1.0]]),
    ],
    "O": [
        (0, [11720.0000000, 1759.0000000, 400.8000000, 113.7000000, 37.0300000,
             13.2700000, 5.0250000, 1.0130000, 0.3023000],
            [[0.0007100, 0.0054700, 0.0278370, 0.1048000, 0.2830620,
              0.4487190, 0.2709520, 0.0154580, -0.0025850],
             [-0.0001600, -0.0012630, -0.0062670, -0.0257160, -0.0709240,
              -0.1654110, -0.1169550, 0.5573680, 0.5727590]]),
        (0, [0.3023000], [[1.0]]),
        (1, [17.7000000, 3.8540000, 1.0460000, 0.2753000],
            [[0.0430180, 0.2289130, 0.5087280, 0.4605310]]),
        (1, [0.2753000], [[1.0]]),
        (2, [1.1850000], [[1.0]]),
    ],
}

_631GS = {
    "H": [
        (0, [18.7311370, 2.8253937, 0.6401217],
            [[0.03349460, 0.23472695, 0.81375733]]),
        (0, [0.1612778], [[1.0]]),
    ],
    "O": [
        (0, [5484.6717000, 825.2349500, 188.0469600, 52.9645000, 16.8975700, 5.7996353],
            [[0.0018311, 0.0139501, 0.0684451, 0.2327143, 0.4701930, 0.3585209]]),
        (0, [15.5396160, 3.5999336, 1.0137618], [[-0.1107775, -0.1480263, 1.1307670]]),
        (1, [15.5396160, 3.5999336, 1.0137618], [[0.0708743, 0.3397528, 0.7271586]]),
        (0, [0.2700058], [[1.0]]),
        (1, [0.2700058], [[1.0]]),
        (2, [0.8000000], [[1.0]]),
    ],
}

BASIS_SETS = {"sto-3g": _STO3G, "cc-pvdz": _CCPVDZ, "6-31g*": _631GS}


def gaussian_int(n: int, alpha):
    """int_0^inf x^n exp(-alpha x^2) dx = Gamma((n+1)/2) / (2 alpha^((n+1)/2))."""
    n1 = (n + 1) * 0.5
    return math.gamma(n1) / (2.0 * np.asarray(alpha, dtype=np.float64) ** n1)


def normalized_coefficients(l: int, exps, coeff_columns) -> np.ndarray:
    """Return coefficients[nprim, nctr] in PySCF's stored normalisation.

    Primitive norm 1/sqrt(gaussian_int(2l+2, 2a)), then each contracted radial
    function r^l sum_p c_p e^{-a_p r^2} is scaled to unit L2 norm with the r^2
    measure.
    """
    es = np.asarray(exps, dtype=np.float64)
    cs = np.asarray(coeff_columns, dtype=np.float64).T  # [nprim, nctr]
    cs = cs / np.sqrt(gaussian_int(2 * l + 2, 2.0 * es))[:, None]
    overlap = gaussian_int(2 * l + 2, es[:, None] + es[None, :])
    norms = np.einsum("pi,pq,qi->i", cs, overlap, cs)
    return cs / np.sqrt(norms)[None, :]


@dataclass
class Molecule:
    """Atoms (symbols, Bohr coordinates) with libcint tables for one basis."""
    symbols: list
    coords: np.ndarray  # [natm, 3] Bohr
    basis: str
    cart: bool
    atm: np.ndarray
    bas: np.ndarray
    env: np.ndarray

    @property
    def natm(self) -> int:
        return len(self.symbols)

    @property
    def nao(self) -> int:
        return nao_of(self.bas, self.cart)

    @property
    def nelectron(self) -> int:
        return int(sum(CHARGES[s] for s in self.symbols))


def nao_of(bas: np.ndarray, cart: bool = False) -> int:
    l = bas[:, ANG_OF].astype(np.int64)
    nc = bas[:, NCTR_OF].astype(np.int64)
    width = (l + 1) * (l + 2) // 2 if cart else 2 * l + 1
    return int(np.sum(width * nc))


def make_tables(symbols, coords_bohr, basis: str = "cc-pvdz", cart: bool = False) -> Molecule:
    """Build libcint-style ATM/BAS/ENV arrays for atoms at ``coords_bohr``."""
    key = basis.lower()
    if key not in BASIS_SETS:
        raise ValueError(f"unknown basis {basis!r}; available: {sorted(BASIS_SETS)}")
    data = BASIS_SETS[key]
    coords = np.ascontiguousarray(coords_bohr, dtype=np.float64).reshape(-1, 3)
    if len(symbols) != len(coords):
        raise ValueError("one coordinate row per atom is required")
    env = [0.0] * PTR_ENV_START
    atm = np.zeros((len(symbols), ATM_SLOTS), dtype=np.int32)
    for i, (sym, xyz) in enumerate(zip(symbols, coords)):
        if sym not in data:
            raise ValueError(f"basis {basis!r} has no data for element {sym!r}")
        atm[i, CHARGE_OF] = CHARGES[sym]
        atm[i, PTR_COORD] = len(env)
        env.extend(float(v) for v in xyz)
        env.append(0.0)  # zeta slot, as PySCF reserves one per atom
    # Shell data are shared by all atoms of one element (PySCF does the same).
    element_ptrs = {}
    for sym in dict.fromkeys(symbols):
        shells = []
        for l, exps, columns in data[sym]:
            coeff = normalized_coefficients(l, exps, columns)
            pexp = len(env)
            env.extend(float(e) for e in exps)
            pcoeff = len(env)
            env.extend(float(c) for c in coeff.T.ravel())  # coeff[ctr*nprim + prim]
            shells.append((l, len(exps), coeff.shape[1], pexp, pcoeff))
        element_ptrs[sym] = shells
    rows = []
    for i, sym in enumerate(symbols):
        for l, nprim, nctr, pexp, pcoeff in element_ptrs[sym]:
            rows.append([i, l, nprim, nctr, 0, pexp, pcoeff, 0])
    bas = np.asarray(rows, dtype=np.int32).reshape(-1, BAS_SLOTS)
    return Molecule(list(symbols), coords, key, cart, atm, bas,
                    np.asarray(env, dtype=np.float64))


def three_s_shell_tables():
    """Hand-built table of three unit-coefficient s shells on three atoms.

    Atoms at (0,0,0), (0,0,1), (0,1,0) with exponents 0.8, 1.1, 1.3 and raw
    coefficient 1 (python/test/test_online_anchor_select.py:19-30).  Each AO is
    exp(-a r^2)/sqrt(4 pi) exactly.
    """
    atm = np.zeros((3, 6), dtype=np.int32)
    atm[:, 1] = [20, 23, 26]
    bas = np.zeros((3, 8), dtype=np.int32)
    bas[:, 0] = [0, 1, 2]
    bas[:, 2:4] = 1
    bas[:, 5] = [29, 31, 33]
    bas[:, 6] = [30, 32, 34]
    env = np.zeros(35)
    env[20:29] = [0, 0, 0, 0, 0, 1.0, 0, 1.0, 0]
    env[29:35] = [0.8, 1, 1.1, 1, 1.3, 1]
    return atm, bas, env
