"""Streaming DE over gene runs of a CSC matrix held in memory-mapped files:
every output equals the oracle (U1 bit for bit, p and FDR within 1e-12,
fold changes within 1e-12 of the oracle's FP64 group means) and the one-shot
hpdex_de_csc on the whole matrix bit for bit, whatever the number of runs and
of runs in flight."""
import sys
from types import SimpleNamespace

import numpy as np
import pandas as pd
import pytest
import scipy.sparse as sp

import oracle
import oracle.de as oracle_de
from fixtures import rel_close
from hpdex_amd import _capi, stream

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _on_disk(tmp_path, m):
    """A CSC whose arrays are read-only memmaps of .npy files."""
    arrs = []
    for name, a in (("data", m.data), ("indices", m.indices.astype(np.int64)),
                    ("indptr", m.indptr.astype(np.int64))):
        f = tmp_path / f"{name}.npy"
        np.save(f, a)
        arrs.append(np.load(f, mmap_mode="r"))
    return arrs


def _case(seed=51, R=6000, C=300, T=9):
    rng = np.random.default_rng(seed)
    X = sp.random(R, C, density=0.1, format="csc", random_state=seed, dtype=np.float64)
    X.data = np.floor(X.data * 30) + 1
    X.data[rng.random(len(X.data)) < 0.2] += 0.5      # a fifth of the values non-integer
    X = X.astype(np.float32).tocsc()
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    return X, gid, T


@pytest.mark.parametrize("workers", [1, 3])
def test_stream_equals_one_shot(tmp_path, workers):
    X, gid, T = _case()
    data, indices, indptr = _on_disk(tmp_path, X)
    runs = stream.plan_gene_runs(indptr, T, stream.run_bytes(30_000, 40, T))
    assert len(runs) >= 5
    say(f"stream: {len(runs)} runs, {workers} in flight")
    got = stream.stream_de(data, indices, indptr, X.shape[0], gid, T,
                           memory_limit_gb=stream.run_bytes(30_000, 40, T) * workers / 2**30,
                           num_workers=workers)
    want = _capi.de_csc(X.data, X.indices.astype(np.int64), X.indptr.astype(np.int64), gid, T,
                        n_rows=X.shape[0])
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)
    # against the oracle: the test (zero_handling min, tie + continuity, as
    # de.py:147-158 calls it), fold changes from its group means, BH FDR
    U1, P, fc, lfc, fdr = got
    oU1, oP = oracle.mwu(X.data, X.indices, X.indptr, gid, T, n_rows=X.shape[0],
                         tie_correction=True, use_continuity=True, zero_handling="min")
    assert np.array_equal(U1, oU1)
    assert rel_close(P, oP, rtol=1e-12)
    means = oracle.group_mean(X.data, X.indices, X.indptr, gid, T + 1, include_zeros=True)
    ofc, olfc = oracle_de.fold_change(means, X.shape[1], T)
    assert rel_close(fc, ofc, rtol=1e-12, atol=1e-15)
    assert rel_close(lfc, olfc, rtol=1e-12, atol=1e-12)
    assert rel_close(fdr, oracle_de.fdr_bh(oP), rtol=1e-12)


def test_stream_driver_target_major(tmp_path):
    X, gid, T = _case(52, R=4000, C=120, T=4)
    data, indices, indptr = _on_disk(tmp_path, X)
    Xd = sp.csc_matrix((data, indices, indptr), shape=X.shape)
    names = np.array(["ctrl"] + [f"KO{i}" for i in range(1, T + 1)] + ["other"])
    labels = np.where(gid < 0, "other", names[np.maximum(gid, 0)])
    obs = pd.DataFrame({"guide": pd.Categorical(labels)})
    adata = SimpleNamespace(obs=obs, X=Xd, var_names=[f"g{i}" for i in range(X.shape[1])])
    say("stream driver")
    df = stream.parallel_differential_expression(adata, "guide", "ctrl",
                                                 groups=[f"KO{i}" for i in range(1, T + 1)],
                                                 memory_limit_gb=0.012, num_workers=2)
    assert list(df.columns) == stream.COLUMNS and len(df) == T * X.shape[1]
    U1, P, fc, lfc, fdr = _capi.de_csc(X.data, X.indices.astype(np.int64),
                                       X.indptr.astype(np.int64), gid, T, n_rows=X.shape[0])
    C = X.shape[1]
    tm = lambda a: a.reshape(C, T).T.reshape(-1)
    assert list(df["target"][:C]) == ["KO1"] * C and list(df["feature"][:2]) == ["g0", "g1"]
    assert np.array_equal(df["p_value"].to_numpy(), tm(P))
    assert np.array_equal(df["u_statistic"].to_numpy(), tm(U1))
    assert np.array_equal(df["fold_change"].to_numpy(), fc)
    assert np.array_equal(df["log2_fold_change"].to_numpy(), lfc)
    assert np.array_equal(df["fdr"].to_numpy(), tm(fdr))
