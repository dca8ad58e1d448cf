"""The CPU oracle against the scipy/numpy golden fixtures (CPU only)."""
import numpy as np
import pytest

import oracle
from fixtures import group_mean_cases, mwu_cases, rel_close, small_config_expected
from hpdex_amd import synthetic

CASES = mwu_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_golden(case):
    U1, P, tie, n1, n2 = oracle.mwu(case["data"], case["indices"], case["indptr"],
                                    case["group_id"], case["n_targets"],
                                    tie_correction=case["tie_correction"],
                                    use_continuity=case["use_continuity"],
                                    zero_handling=case["zero_handling"],
                                    alternative=case["alternative"], method=case["method"],
                                    threads=2, with_stats=True)
    assert np.array_equal(U1, case["U1"]), "U1 must be bit-exact"
    assert np.array_equal(tie, case["tie"]), "tie term must be bit-exact"
    assert rel_close(P, case["P"], rtol=1e-12), np.max(np.abs(P - case["P"]))


@pytest.mark.parametrize("k", range(4))
def test_oracle_group_mean_golden(k):
    c = group_mean_cases()[k]
    for kahan in (False, True):
        out = oracle.group_mean(c["data"], c["indices"], c["indptr"], c["group_id"],
                                c["n_groups"], include_zeros=c["include_zeros"], use_kahan=kahan)
        assert rel_close(out, c["expected"], rtol=1e-12, atol=1e-12)


def test_oracle_small_config():
    w = synthetic.small_case()
    exp = small_config_expected()
    U1, P, tie, _, _ = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], 4,
                                  tie_correction=True, use_continuity=True, zero_handling="min",
                                  with_stats=True)
    assert np.array_equal(U1, exp["U1"])
    assert np.array_equal(tie, exp["tie"])
    assert rel_close(P, exp["P"], rtol=1e-12)
    gm = oracle.group_mean(w["data"], w["indices"], w["indptr"], w["group_id"], 5)
    assert rel_close(gm, exp["group_mean"], rtol=1e-12, atol=1e-15)
