"""The DE post-processing oracle (oracle/de.py) pinned on the CPU: BH FDR
equals scipy.stats.false_discovery_control (the routine de.py calls) bit for
bit, including ties, boundaries and its failure; fold change follows the clip
rules and layout of de.py:126-142 on hand-checked values."""
import numpy as np
import pytest
import scipy.stats

from oracle import de as ode


@pytest.mark.parametrize("case", ["uniform", "ties", "boundaries", "tiny", "descending"])
def test_fdr_matches_scipy(case):
    rng = np.random.default_rng(7)
    if case == "uniform":
        p = rng.random(5000)
    elif case == "ties":
        p = np.round(rng.random(3000), 2)
    elif case == "boundaries":
        p = np.concatenate([np.zeros(5), np.ones(7), rng.random(50), [1e-300, 5e-324]])
    elif case == "tiny":
        p = np.array([0.3, 0.01])
    else:
        p = np.linspace(1.0, 0.0, 777)
    np.testing.assert_array_equal(ode.fdr_bh(p), scipy.stats.false_discovery_control(p, method="bh"))


def test_fdr_single_and_empty():
    assert ode.fdr_bh(np.array([0.7]))[0] == 0.7
    assert ode.fdr_bh(np.array([])).size == 0


def test_fdr_invalid_gives_nan_where_scipy_raises():
    p = np.array([0.1, np.nan, 0.5])
    with pytest.raises(ValueError):
        scipy.stats.false_discovery_control(p, method="bh")
    assert np.all(np.isnan(ode.fdr_bh(p)))
    assert np.all(np.isnan(ode.fdr_bh(np.array([0.2, 1.5]))))


def test_fdr_planted_off_by_one_is_caught():
    """Scaling the k-th value by m/k+1 instead of m/k would change the result."""
    p = np.random.default_rng(3).random(100)
    order = np.argsort(p)
    wrong = p[order] * (100 / np.arange(2, 102))
    wrong = np.minimum.accumulate(wrong[::-1])[::-1]
    out = np.empty(100)
    out[order] = wrong
    assert not np.array_equal(np.clip(out, 0, 1), ode.fdr_bh(p))


def test_fold_change_rules_and_layout():
    # 2 genes x (reference + 3 targets), group_mean layout c*G + g
    means = np.array([[2.0, 4.0, 0.0, 1.0],      # gene 0
                      [0.0, 3.0, 1e-12, 0.5]])   # gene 1: reference ~ 0
    fc, lfc = ode.fold_change(means.reshape(-1), n_cols=2, n_targets=3, clip_value=20.0)
    # target-major: t*C + c
    expect = np.array([[2.0, 20.0],              # target 1: 4/2, ref 0 -> clip
                       [1 / 20.0, 1 / 20.0],     # target 2: tar < 1e-10 -> 1/clip (wins)
                       [0.5, 20.0]])             # target 3: 1/2, ref 0 -> clip
    np.testing.assert_array_equal(fc, expect.reshape(-1))
    np.testing.assert_array_equal(lfc, np.log2(expect.reshape(-1)))
