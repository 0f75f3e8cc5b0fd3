"""The oracle against the reference's own accuracy logs (script/a100/*.dat, tests/golden/a100_kat.json).

Each log row holds the max relative error (hex) and the number of entries that differ from
RN(exact) for one size of a sweep of testing_ozblas_d{gemm,dot,gemv}.  Regenerating the harness
inputs and replaying the sweep through the oracle must give identical rows: this pins the split
(slice boundaries and counts), the Fsum order, NearSum and the GEMM/GEMV/DOT drivers bit for bit.
"""
import json
import os

import pytest

import kat_harness as K

_FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "a100_kat.json")))
CASES = _FIX["cases"]


@pytest.mark.parametrize("case", CASES, ids=[c["dat"] for c in CASES])
def test_reference_log(case, oracle):
    th = K.harness_opts(case["opts"])
    impl = K.OracleImpl(K.handle_flags(th))
    if th["fastMode"]:
        # the fastmode=1 logs were produced with the legacy level rule (see ozoracle.h)
        impl.h.set(fastModeRule=1)
    rows = K.RUNNERS[case["routine"]](case["opts"], impl)
    assert len(rows) == len(case["rows"])
    assert K.compare(case, rows) == []


def test_fastmode_rule_matters(oracle):
    """The shipped fastMode rule and the legacy one give different bits on a logged case, so the
    log comparison above discriminates between them."""
    case = next(c for c in CASES if c["dat"] == "_dgemm_tt_sum0_fast1_bgemm0_sq_s.dat")
    th = K.harness_opts(case["opts"])
    impl = K.OracleImpl(K.handle_flags(th))
    rows = K.RUNNERS["gemm"](case["opts"], impl)
    assert K.compare(case, rows) != []
