"""Pins of the numba-pipeline oracle (oracle/pden.py) against what fixes it
independently: scipy.stats.mannwhitneyu (asymptotic and exact), a brute-force
enumeration of the Mann-Whitney null, closed-form means and fold changes, and
scipy's BH.  CPU only."""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse
import scipy.stats as ss

from oracle import pden as op


def _close(a, b, rtol):
    a, b = np.asarray(a, float), np.asarray(b, float)
    assert np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(b), 1e-300)), (a, b)


@pytest.mark.parametrize("cc", [False, True])
def test_merge_and_asymptotic_p_match_scipy(cc):
    rng = np.random.default_rng(11)
    for _ in range(6):
        nr, nt = int(rng.integers(5, 60)), int(rng.integers(3, 40))
        ref = rng.poisson(1.0, (9, nr)).astype(np.float64)   # zeros and ties
        tar = rng.poisson(1.4, (9, nt)).astype(np.float64)
        p, U = op.rank_sum_float(np.sort(ref, 1), np.sort(tar, 1), True, cc, True)
        for b in range(9):
            r = ss.mannwhitneyu(tar[b], ref[b], use_continuity=cc, alternative="two-sided",
                                method="asymptotic")
            assert r.statistic == U[b]   # the TARGET group's U
            _close(p[b], r.pvalue, 1e-12)


def test_asymptotic_without_tie_correction_is_the_plain_normal_tail():
    rng = np.random.default_rng(12)
    ref = rng.poisson(0.8, (5, 40)).astype(float)
    tar = rng.poisson(1.5, (5, 25)).astype(float)
    p, U = op.rank_sum_float(np.sort(ref, 1), np.sort(tar, 1), False, True, True)
    n1, n2 = 25, 40
    for b in range(5):
        z = (abs(U[b] - n1 * n2 / 2) - 0.5) / math.sqrt(n1 * n2 * (n1 + n2 + 1) / 12)
        _close(p[b], 2 * ss.norm.sf(z), 1e-12)


@pytest.mark.parametrize("n1,n2", [(1, 1), (2, 3), (3, 3), (4, 2), (3, 5), (5, 4)])
def test_exact_table_is_the_enumerated_null(n1, n2):
    N = n1 + n2
    cnt = np.zeros(n1 * n2 + 1, dtype=np.int64)
    for pos in itertools.combinations(range(1, N + 1), n1):
        cnt[sum(pos) - n1 * (n1 + 1) // 2] += 1
    sf = np.cumsum(cnt[::-1])[::-1] / cnt.sum()
    assert np.array_equal(op.exact_tail_table(n1, n2), sf)


def test_exact_p_matches_scipy_exact():
    rng = np.random.default_rng(13)
    for _ in range(12):
        nr, nt = int(rng.integers(2, 9)), int(rng.integers(2, 9))
        ref, tar = rng.normal(size=(3, nr)), rng.normal(size=(3, nt))
        p, U = op.rank_sum_float(np.sort(ref, 1), np.sort(tar, 1), True, True, None)
        for b in range(3):
            r = ss.mannwhitneyu(tar[b], ref[b], alternative="two-sided", method="exact")
            assert r.statistic == U[b]
            _close(p[b], r.pvalue, 1e-12)


def test_exact_index_rounds_half_to_even():
    # forced exact with ties: U is a half-integer and np.rint picks the even
    # neighbour (pykernel_.py:283-287)
    ref = np.array([[0.0, 1.0, 2.0]])
    tar = np.array([[1.0, 3.0]])      # ranks 2.5 and 5: U_tar = 4.5, U_ref = 1.5
    p, U = op.rank_sum_float(ref, tar, True, True, False)
    assert U[0] == 4.5
    sf = op.exact_tail_table(2, 3)
    assert p[0] == min(1.0, 2 * sf[4])   # rint(4.5) = 4, not 5
    ref = np.array([[0.0, 1.0, 2.0, 3.0]])
    tar = np.array([[1.0, 4.0]])      # ranks 2.5 and 6: U_tar = 5.5
    p, U = op.rank_sum_float(ref, tar, True, True, False)
    sf = op.exact_tail_table(2, 4)
    assert U[0] == 5.5 and sf[6] != sf[5]
    assert p[0] == min(1.0, 2 * sf[6])   # rint(5.5) = 6: rounding down would give sf[5]


def test_chunk_method_choice():
    """One worker: one task per group; 8 target cells and no tie anywhere ->
    exact; a tie in any gene of the task -> asymptotic for every gene."""
    rng = np.random.default_rng(14)
    R, C = 20, 6
    X = rng.normal(size=(R, C)) + 5.0
    labels = np.array(["ref"] * 12 + ["a"] * 8)
    df = op.pden(labels, X, [f"g{i}" for i in range(C)], "ref")
    for c in range(C):
        r = ss.mannwhitneyu(X[12:, c], X[:12, c], alternative="two-sided", method="exact")
        _close(df.p_value[c], r.pvalue, 1e-12)
    X[0, 3] = X[1, 3]                    # one tie in gene 3
    df = op.pden(labels, X, [f"g{i}" for i in range(C)], "ref")
    for c in range(C):
        r = ss.mannwhitneyu(X[12:, c], X[:12, c], alternative="two-sided", method="asymptotic")
        _close(df.p_value[c], r.pvalue, 1e-12)


def test_hist_fallback_forces_asymptotic():
    rng = np.random.default_rng(15)
    R, C = 14, 3
    X = rng.permutation(np.arange(R * C)).reshape(R, C).astype(np.int64) * 1000  # no ties
    labels = np.array(["ref"] * 8 + ["a"] * 6)
    kw = dict(use_asymptotic=False, prefer_hist_if_int=True, num_workers=2)
    big = op.pden(labels, X, ["x", "y", "z"], "ref", max_bins=100, **kw)
    small = op.pden(labels, X, ["x", "y", "z"], "ref", max_bins=10 ** 9, **kw)
    for c in range(C):
        ra = ss.mannwhitneyu(X[8:, c], X[:8, c], alternative="two-sided", method="asymptotic")
        re = ss.mannwhitneyu(X[8:, c], X[:8, c], alternative="two-sided", method="exact")
        _close(big.p_value[c], ra.pvalue, 1e-12)
        _close(small.p_value[c], re.pvalue, 1e-12)


def test_fold_change_closed_forms():
    # integer counts: every mean is exact
    X = np.array([[0, 2, 0, 4], [0, 4, 0, 0], [3, 0, 0, 1], [1, 0, 0, 1]], dtype=np.float32)
    labels = np.array(["ref", "ref", "a", "a"])
    df = op.pden(labels, X, list("wxyz"), "ref", min_samples=1)
    # ref means 0, 3, 0, 2; target means 2, 0, 0, 1
    assert df.fold_change.tolist() == [20.0, np.float32(1 / 20.0), 20.0, 0.5]
    assert df.log2_fold_change[3] == -1.0
    assert df.fold_change.dtype == np.float64


def test_fdr_is_scipy_bh_over_all_groups():
    rng = np.random.default_rng(16)
    X = scipy.sparse.random(60, 30, density=0.4, random_state=3, format="csr",
                            data_rvs=lambda n: rng.integers(1, 9, n)).astype(np.float32)
    labels = np.array(["ref"] * 20 + ["a"] * 15 + ["b"] * 25)
    df = op.pden(labels, X, [f"g{i}" for i in range(30)], "ref", groups=["a", "b"])
    _close(df.fdr, ss.false_discovery_control(df.p_value, method="bh"), 1e-15)
    # the target U and the reference U of hpdex.backend.mannwhitneyu add to n1 n2
    Xd = X.toarray()
    for c in range(3):
        r = ss.mannwhitneyu(Xd[:20, c], Xd[20:35, c])
        assert df.u_statistic[c] + r.statistic == 20 * 15
