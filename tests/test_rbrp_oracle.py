"""CPU: the oracle's RBRP and ID (cpp/src/anchor_select_rbrp.cpp:135-318) on
the reference's own RBRP checks (python/test/test_rbrp_id.py:17-59)."""
import oracle
import rbrp_cases as cases


def test_direct_fit_at_pivots():
    cases.check_direct_fit_at_pivots(oracle.rbrp_ao)


def test_absolute_weighted_selection():
    cases.check_absolute_weighted_selection(oracle.rbrp_ao)


def test_tolerance_stop():
    cases.check_tolerance_stop(oracle.rbrp_ao)
