"""DE post-processing on the GPU (hpdex_fdr_bh, hpdex_de_csc and the
parallel_differential_expression driver) against the oracle: U1 bit-exact,
p / fold change / log2 fold change / FDR within 1e-12 relative."""
import sys
from types import SimpleNamespace

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp
import scipy.stats

import oracle
from fixtures import rel_close
from hpdex_amd import _capi, synthetic
from oracle import de as ode

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("case", ["uniform", "ties", "boundaries", "single"])
def test_fdr_bh_matches_scipy(case):
    say(f"fdr {case}")
    rng = np.random.default_rng(11)
    p = {"uniform": rng.random(200_000),
         "ties": np.round(rng.random(50_000), 3),
         "boundaries": np.concatenate([np.zeros(9), np.ones(9), rng.random(1000), [5e-324]]),
         "single": np.array([0.42])}[case]
    got = _capi.fdr_bh(p)
    want = ode.fdr_bh(p)
    np.testing.assert_array_equal(want, scipy.stats.false_discovery_control(p, method="bh"))
    assert rel_close(got, want, rtol=RTOL)


def test_fdr_bh_invalid_is_nan():
    say("fdr nan")
    assert np.all(np.isnan(_capi.fdr_bh(np.array([0.1, np.nan, 0.3]))))


def test_fdr_bh_large_properties():
    """2.4M p-values (the tests of the 150,000 x 8,000 workload): FDR >= p,
    monotone in p, equal to the oracle."""
    say("fdr large")
    p = np.random.default_rng(12).beta(0.3, 1.0, 2_400_000)
    got = _capi.fdr_bh(p)
    assert np.all(got >= p - 1e-300)
    order = np.argsort(p, kind="stable")
    assert np.all(np.diff(got[order]) >= 0)
    assert rel_close(got, ode.fdr_bh(p), rtol=RTOL)


def _oracle_de(w, clip=20.0, tie=True, cc=True):
    T, C = w["n_targets"], len(w["indptr"]) - 1
    means = oracle.group_mean(w["data"], w["indices"], w["indptr"], w["group_id"], T + 1,
                              include_zeros=True)
    fc, lfc = ode.fold_change(means, C, T, clip)
    U1, P = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], T,
                       n_rows=w["n_rows"], tie_correction=tie, use_continuity=cc,
                       zero_handling="min")
    return U1, P, fc, lfc, ode.fdr_bh(P)


def _check(got, want):
    U1g, Pg, fcg, lfcg, fdrg = got
    U1o, Po, fco, lfco, fdro = want
    assert np.array_equal(U1g, U1o)
    assert rel_close(Pg, Po, rtol=RTOL)
    assert rel_close(fcg, fco, rtol=RTOL)
    assert rel_close(lfcg, lfco, rtol=RTOL, atol=1e-15)
    assert rel_close(fdrg, fdro, rtol=RTOL)


def test_de_csc_small_workload():
    say("de small")
    w = synthetic.small_case()
    got = _capi.de_csc(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                       n_rows=w["n_rows"])
    _check(got, _oracle_de(w))


def test_de_csc_clip_rules_and_float_values():
    """Genes absent from the reference or from a target hit the clip rules;
    float values take the sort path."""
    say("de clip")
    rng = np.random.default_rng(13)
    R, C, T = 3000, 30, 5
    X = rng.random((R, C)) * (rng.random((R, C)) < 0.3)
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    X[gid == 0, :5] = 0.0          # reference mean 0 in genes 0-4
    X[gid == 2, 5:9] = 0.0         # target 2 mean 0 in genes 5-8
    m = sp.csc_matrix(X.astype(np.float32))
    w = dict(data=m.data, indices=m.indices.astype(np.int64), indptr=m.indptr.astype(np.int64),
             group_id=gid, n_targets=T, n_rows=R)
    got = _capi.de_csc(w["data"], w["indices"], w["indptr"], gid, T, n_rows=R, clip_value=16.0)
    want = _oracle_de(w, clip=16.0)
    _check(got, want)
    fc = got[2].reshape(T, C)
    assert np.all(fc[:, :5] == 16.0)          # reference mean 0 -> clip
    assert np.all(fc[1, 5:9] == 1 / 16.0)     # target 2 mean 0 -> 1/clip


def test_parallel_differential_expression_driver():
    """The driver on an AnnData-shaped object: same DataFrame as the
    reference's steps computed with the oracle (labels in de.py's layout)."""
    say("de driver")
    from hpdex_amd.de import parallel_differential_expression
    rng = np.random.default_rng(14)
    R, C = 2500, 40
    X = sp.random(R, C, density=0.15, format="csr", random_state=15, dtype=np.float32)
    X.data = np.floor(X.data * 20) + 1
    labels = rng.choice(["ctrl", "KO1", "KO2", "KO3", "tiny"], R, p=[0.4, 0.2, 0.2, 0.2, 0.0])
    labels[:1] = "tiny"                               # one cell: dropped by min_samples
    obs = pd.DataFrame({"guide": pd.Categorical(labels)})
    adata = SimpleNamespace(obs=obs, X=X, var_names=[f"g{i}" for i in range(C)], n_vars=C)
    df = parallel_differential_expression(adata, "guide", "ctrl")
    groups = [g for g in obs["guide"].unique().tolist() if g not in ("ctrl", "tiny")]
    assert list(df.columns) == ["target", "feature", "p_value", "u_statistic", "fold_change",
                                "log2_fold_change", "fdr"]
    assert len(df) == C * len(groups)
    assert list(df["target"][:C]) == [groups[0]] * C
    gid = np.full(R, -1, np.int32)
    gid[labels == "ctrl"] = 0
    for i, g in enumerate(groups):
        gid[labels == g] = i + 1
    m = X.tocsc()
    w = dict(data=m.data, indices=m.indices.astype(np.int64), indptr=m.indptr.astype(np.int64),
             group_id=gid, n_targets=len(groups), n_rows=R)
    U1, P, fc, lfc, fdr = _oracle_de(w)
    _check((df["u_statistic"].to_numpy(), df["p_value"].to_numpy(), df["fold_change"].to_numpy(),
            df["log2_fold_change"].to_numpy(), df["fdr"].to_numpy()), (U1, P, fc, lfc, fdr))
