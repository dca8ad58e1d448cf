"""hpdex.pden on the device (hpdex_amd.backend.de_ -> hpdex_pden_matrix)
against the oracle's restatement of the numba pipeline (oracle/pden.py), on
the same seeded inputs.  Tolerances: target U bit-exact (half-integer sums);
p within 1e-12 relative (the asymptotic p's erfc is the device's, the exact
p is the same count sums and division); fold change bit-exact (numpy's
float32 / float64 mean reproduced operation by operation); log2 fold change
within 2 float32 ulps for a float32 X (numpy's float32 log2 is not
correctly rounded) and 1e-15 otherwise; FDR within 1e-12."""
from types import SimpleNamespace

import numpy as np
import pandas as pd
import pytest
import scipy.sparse

pytestmark = pytest.mark.gpu


def _adata(X, labels, key="grp"):
    R, C = X.shape
    return SimpleNamespace(obs=pd.DataFrame({key: pd.Categorical(labels)}), X=X,
                           var=pd.DataFrame(index=[f"g{i}" for i in range(C)]), n_vars=C,
                           n_obs=R, shape=(R, C))


def _run_both(X, labels, reference, **kw):
    from hpdex_amd.backend import parallel_differential_expression_numba as pden
    from oracle import pden as op
    ad = _adata(X, labels)
    got = pden(ad, "grp", reference, **kw)
    exp = op.pden(np.asarray(labels), X, ad.var.index.values, reference, **kw)
    return got, exp


def _check(got, exp, f32):
    assert list(got.columns) == list(exp.columns)
    assert got.target.tolist() == exp.target.tolist()
    assert got.feature.tolist() == exp.feature.tolist()
    assert np.array_equal(got.u_statistic.values, exp.u_statistic.values)
    p, q = got.p_value.values, exp.p_value.values
    assert np.all(np.abs(p - q) <= 1e-12 * np.maximum(q, 1e-300)), np.max(np.abs(p - q))
    assert np.array_equal(got.fold_change.values, exp.fold_change.values)
    a, b = got.log2_fold_change.values, exp.log2_fold_change.values
    if f32:
        tol = 2 * np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    else:
        tol = 1e-15 * np.abs(b)
    assert np.all(np.abs(a - b) <= tol)
    f, g = got.fdr.values, exp.fdr.values
    assert np.all(np.abs(f - g) <= 1e-12 * np.maximum(g, 1e-300))


def test_pden_float32_sparse_auto():
    rng = np.random.default_rng(31)
    R, C = 900, 300
    X = scipy.sparse.random(R, C, density=0.15, random_state=5, format="csr", dtype=np.float32)
    X.data = np.log1p(rng.lognormal(0.0, 1.0, X.nnz)).astype(np.float32)
    labels = rng.choice(["ctrl", "A", "B", "C", "D"], size=R, p=[0.4, 0.2, 0.2, 0.15, 0.05])
    print("pden float32 sparse", R, C, flush=True)
    got, exp = _run_both(X, labels, "ctrl")
    _check(got, exp, True)


def test_pden_int_hist_chunks_forced_exact():
    """Integer counts with the histogram kernel over two gene chunks (8,200
    genes, 2 workers -> 4,100 per task): chunk 1 passes max_bins and falls
    back to the asymptotic p, chunk 0 takes the forced exact p with ties
    (half-integer U, rint)."""
    rng = np.random.default_rng(32)
    R, C = 26, 8200
    X = rng.poisson(1.2, (R, C)).astype(np.int32)
    X[:, 4100:] += (rng.random((R, C - 4100)) < 0.01) * 500   # range > max_bins in chunk 1
    labels = np.array(["ref"] * 10 + ["a"] * 6 + ["b"] * 4 + ["c"] * 6)
    print("pden int hist", R, C, flush=True)
    got, exp = _run_both(X, labels, "ref", use_asymptotic=False, prefer_hist_if_int=True,
                         num_workers=2, max_bins=100, groups=["a", "b", "c"])
    _check(got, exp, False)
    # both methods did occur
    assert np.any(got.u_statistic.values % 1 == 0.5)


def test_pden_small_groups_exact_auto_float64():
    rng = np.random.default_rng(33)
    R, C = 40, 64
    X = rng.normal(size=(R, C)) + 3.0       # no zeros, no ties
    labels = np.array(["ref"] * 12 + ["s3"] * 3 + ["s5"] * 5 + ["s8"] * 8 + ["big"] * 12)
    X[12, 7] = X[0, 7]                      # one tie: group s3 goes asymptotic
    print("pden small exact float64", R, C, flush=True)
    got, exp = _run_both(X, labels, "ref", groups=["s3", "s5", "s8", "big"])
    _check(got, exp, False)


def test_pden_min_samples_one_and_continuity_off():
    rng = np.random.default_rng(34)
    R, C = 9, 20
    X = rng.poisson(2.0, (R, C)).astype(np.float32)
    labels = np.array(["ref"] + ["one"] + ["x"] * 7)
    print("pden min_samples 1", flush=True)
    got, exp = _run_both(X, labels, "ref", min_samples=1, continuity_correction=False,
                         tie_correction=False, groups=["one", "x"])
    _check(got, exp, True)


def test_pden_messages():
    from hpdex_amd.backend import parallel_differential_expression_numba as pden
    X = np.ones((6, 3), dtype=np.float32)
    ad = _adata(X, ["r", "r", "a", "a", "b", "b"])
    with pytest.raises(ValueError, match="reference `zz` not found in `grp`"):
        pden(ad, "grp", "zz")
    with pytest.raises(ValueError, match="No target groups meet minimum sample size requirement"):
        pden(ad, "grp", "r", min_samples=3)
    with pytest.raises(ValueError, match="Unsupported statistical test method: t-test"):
        pden(ad, "grp", "r", metric="t-test")
    with pytest.raises(TypeError, match="float16 is not supported"):
        pden(_adata(X.astype(np.float16), ["r", "r", "a", "a", "b", "b"]), "grp", "r")
