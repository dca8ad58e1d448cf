"""Pins of the oracle beyond the fixtures: brute force on tiny inputs, the
exact null against enumeration and scipy, invariants, error strings."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.stats as st

import oracle


def brute_midrank(a, b):
    """U1 and tie term by the definition: midranks of the pooled sample."""
    pooled = np.concatenate([a, b])
    ranks = st.rankdata(pooled)          # average ranks
    R1 = ranks[:len(a)].sum()
    U1 = R1 - len(a) * (len(a) + 1) / 2
    _, cnt = np.unique(pooled, return_counts=True)
    return U1, float(np.sum(cnt.astype(float) ** 3 - cnt))


@pytest.mark.parametrize("seed", range(25))
def test_brute_force_none_mode(seed):
    rng = np.random.default_rng(seed)
    R, C, T = 24, 3, 2
    dense = rng.integers(-2, 3, (R, C)).astype(np.float64)
    gid = np.repeat([0, 1, 2], 8).astype(np.int32)
    # keep stored values (including explicit zeros) via coo->csc with explicit entries
    rows, cols = np.nonzero(np.ones_like(dense))
    M = sp.csc_matrix((dense[rows, cols], (rows, cols)), shape=(R, C))
    M.sort_indices()
    U1, P, tie, n1, n2 = oracle.mwu(M.data, M.indices, M.indptr, gid, T, zero_handling="none",
                                    with_stats=True)
    for c in range(C):
        for g in (1, 2):
            a, b = dense[gid == 0, c], dense[gid == g, c]
            u, t = brute_midrank(a, b)
            i = c * T + g - 1
            assert U1[i] == u and tie[i] == t
            assert U1[i] + (n1[i] * n2[i] - U1[i]) == n1[i] * n2[i]
            assert 0 <= U1[i] <= n1[i] * n2[i] and (2 * U1[i]) == int(2 * U1[i])


@pytest.mark.parametrize("n1,n2", [(1, 1), (2, 1), (1, 5), (3, 4), (4, 4), (5, 3), (6, 2), (5, 5)])
def test_exact_counts_enumeration(n1, n2):
    c = oracle.exact_counts(n1, n2)
    bf = np.zeros(n1 * n2 + 1)
    for comb in itertools.combinations(range(1, n1 + n2 + 1), n1):
        bf[sum(comb) - n1 * (n1 + 1) // 2] += 1
    assert np.array_equal(c, bf)
    assert c.sum() == float(np.round(np.exp(np.sum(np.log(np.arange(1, n1 + n2 + 1))) -
                                            np.sum(np.log(np.arange(1, n1 + 1))) -
                                            np.sum(np.log(np.arange(1, n2 + 1))))))


@pytest.mark.parametrize("n1,n2", [(8, 5), (12, 12), (25, 7), (60, 3), (60, 25), (30, 19)])
def test_exact_p_vs_scipy(n1, n2):
    rng = np.random.default_rng(n1 * 100 + n2)
    for alt in ("two-sided", "less", "greater"):
        for _ in range(4):
            x = rng.normal(size=n1) + rng.normal(0, 0.7)
            y = rng.normal(size=n2)
            r = st.mannwhitneyu(x, y, method="exact", alternative=alt)
            p = oracle.p_exact(r.statistic, n1, n2, alt)
            assert abs(p - r.pvalue) <= 1e-12 * max(p, r.pvalue) + 1e-300, (alt, p, r.pvalue)


def test_exact_counts_symmetric():
    for n1, n2 in [(7, 9), (13, 4)]:
        c = oracle.exact_counts(n1, n2)
        assert np.array_equal(c, c[::-1])
        assert np.array_equal(c, oracle.exact_counts(n2, n1))


def test_asymptotic_known_values():
    # sigma = 0 (every value tied) gives p = 1
    assert oracle.p_asymptotic(8.0, 4, 4, 8 ** 3 - 8, 0.0) == 1.0
    # symmetric statistic, no correction: z = 0 -> p = 1 two-sided, 0.5 one-sided
    assert oracle.p_asymptotic(8.0, 4, 4, 0.0, 0.0) == 1.0
    assert oracle.p_asymptotic(8.0, 4, 4, 0.0, 0.0, "greater") == 0.5
    # against scipy.stats.norm on a closed-form z
    n1, n2, U = 30, 40, 800.0
    sd = np.sqrt(n1 * n2 * (n1 + n2 + 1) / 12)
    z = (U - n1 * n2 / 2 - 0.5) / sd
    assert abs(oracle.p_asymptotic(U, n1, n2, 0.0, 0.5, "greater") - st.norm.sf(z)) < 1e-15


def _tiny(gid_sizes, T, zero_handling="min", col_vals=None):
    R = sum(gid_sizes)
    gid = np.repeat(np.arange(len(gid_sizes)), gid_sizes).astype(np.int32)
    dense = np.arange(1, R + 1, dtype=np.float32)[:, None].repeat(2, 1) if col_vals is None else col_vals
    M = sp.csc_matrix(dense)
    return M, gid


def test_error_messages():
    M, gid = _tiny([1, 3], 1)
    with pytest.raises(oracle.OracleError, match=r"^Sample too small for reference at column 0$"):
        oracle.mwu(M.data, M.indices, M.indptr, gid, 1, zero_handling="min")
    M, gid = _tiny([3, 3, 1], 2)
    with pytest.raises(oracle.OracleError, match=r"^Sample too small for group 2 at column 0$"):
        oracle.mwu(M.data, M.indices, M.indptr, gid, 2, zero_handling="mix")
    with pytest.raises(oracle.OracleError, match=r"n_targets must be >= 1"):
        oracle.mwu(M.data, M.indices, M.indptr, gid, 0)
    # none mode: sizes per column; column 1 of the target has a single stored value
    dense = np.array([[1, 1], [2, 2], [3, 0], [4, 5], [5, 0]], dtype=np.float32)
    gid = np.array([0, 0, 1, 1, 1], dtype=np.int32)
    dense[2, 1] = 0
    dense[4, 1] = 0
    M = sp.csc_matrix(dense)
    with pytest.raises(oracle.OracleError, match=r"^Sample too small for group 1 at column 1$"):
        oracle.mwu(M.data, M.indices, M.indptr, gid, 1, zero_handling="none")


def test_planted_mistakes_are_caught():
    """A wrong zero position or a swapped operand changes U1 on a fixture-sized input."""
    rng = np.random.default_rng(3)
    a = np.array([-1.0, 0, 0, 2, 3]); b = np.array([0.0, 1, -2, 5])
    u_ok, _ = brute_midrank(np.where(a == 0, -np.inf, a), np.where(b == 0, -np.inf, b))
    u_mix, _ = brute_midrank(a, b)
    u_swap, _ = brute_midrank(np.where(b == 0, -np.inf, b), np.where(a == 0, -np.inf, a))
    assert u_ok != u_mix and u_ok != u_swap
    dense = np.zeros((9, 1)); dense[:5, 0] = a; dense[5:, 0] = b
    gid = np.array([0] * 5 + [1] * 4, dtype=np.int32)
    M = sp.csc_matrix(dense)
    U1, _ = oracle.mwu(M.data, M.indices, M.indptr, gid, 1, zero_handling="min")
    assert U1[0] == u_ok
    del rng
