"""Omega (cpp/src/anchor_select_randomized_id.cpp:269-279) is host code in the
product: its libstdc++ stream must equal the oracle's bit for bit."""
import numpy as np

import oracle
from mp_isdf_amd import capi


def test_product_omega_equals_oracle_bitwise():
    # The larger cases span several parallel chunks of 65,536 polar attempts
    # (the last about six), copied out in parallel, and end on odd counts.
    for d, ell, seed in ((7, 13, 7), (24, 30, 31), (3, 1, 0), (96, 1101, 7), (129, 517, 2**63 + 5),
                         (256, 2101, 11)):
        u = np.empty((d, ell))
        v = np.empty((d, ell))
        capi.call("mpisdf_omega", d, ell, seed, u, v)
        ou, ov = oracle.omega(d, ell, seed)
        np.testing.assert_array_equal(u, ou)
        np.testing.assert_array_equal(v, ov)
