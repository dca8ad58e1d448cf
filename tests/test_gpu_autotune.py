"""The autotuner on the device: every candidate layout of a job (histogram /
sort cutover, LDS per histogram workgroup, workgroup sizes 256/512/1024)
gives the same U1 and p bit for bit, the fastest is applied, and the job run
under it matches the oracle."""
import sys

import numpy as np
import pytest

import oracle
from hpdex_amd import _capi, autotune, synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def restore_tuning():
    t = autotune.current()
    yield
    autotune.apply(t)


def _mixed_workload(seed=5, R=12000, C=24, T=30):
    # integer counts in 16 columns (histogram path when the cutover admits
    # them), continuous values in 8 (always the sort path)
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, 0.3))
    data = (rng.negative_binomial(0.8, 0.05, len(rows)) + 1).astype(np.float32)
    for c in range(16, C):
        s, e = indptr[c], indptr[c + 1]
        data[s:e] = rng.normal(0, 1, e - s).astype(np.float32)
    gid = np.full(R, -1, dtype=np.int32)
    perm = rng.permutation(R)
    gid[perm[:3000]] = 0
    gid[perm[3000:]] = rng.integers(1, T + 1, R - 3000)
    return dict(data=data, indices=rows, indptr=indptr, group_id=gid, n_targets=T, n_rows=R)


@pytest.mark.parametrize("mode", ["min", "mix"])
def test_autotune_candidates_agree_and_match_oracle(mode, restore_tuning):
    w = _mixed_workload()
    zh = {"min": 1, "mix": 3}[mode]
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=zh)
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    try:
        say(f"autotune mode={mode}: {len(autotune.candidates())} candidates")
        res = autotune.autotune(m, w["group_id"], w["n_targets"], opts, iters=3, log=say)
        assert len(res["trials"]) == len(autotune.candidates())
        assert all(t["ms"] > 0 for t in res["trials"])
        sort_only = [t for t in res["trials"] if t["hist_max_bins"] == 0]
        assert sort_only and sort_only[0]["count_columns"] == 0
        assert max(t["count_columns"] for t in res["trials"]) == 16
        assert autotune.current() == res["best"]
        p = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1)
        try:
            p.run(1)
            U1, P = p.results()
        finally:
            p.close()
    finally:
        m.close()
    o = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                   n_rows=w["n_rows"], tie_correction=True, use_continuity=True,
                   zero_handling=mode, alternative="two-sided", method="asymptotic",
                   with_stats=True)
    np.testing.assert_array_equal(U1, o[0])
    np.testing.assert_allclose(P, o[1], rtol=1e-12, atol=0)


def test_autotune_sort_cutover_with_a_reference_beyond_the_chip(restore_tuning, tmp_path):
    # an integer column whose ~12,000 reference entries span values 1..3000
    # (about 2,900 distinct): under a 2,048-bin cutover it takes the sort path, where its reference
    # slice passes the 8,192 keys sorted on chip and is sorted in HBM; under
    # 8,192 bins it takes the histogram kernel.  Both candidates run and agree
    # bit for bit; the saved choice is what the library holds after a load.
    rng = np.random.default_rng(9)
    R, T = 40000, 4
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.6, 0.1]))
    data = rng.integers(1, 3001, len(rows)).astype(np.float32)
    gid = np.full(R, -1, dtype=np.int32)
    perm = rng.permutation(R)
    gid[perm[:20000]] = 0
    gid[perm[20000:]] = rng.integers(1, T + 1, R - 20000)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    try:
        say("autotune across the sort cutover")
        res = autotune.autotune(m, gid, T, opts, space={"hist_max_bins": (2048, 8192)},
                                iters=2, log=say)
        sort, hist = res["trials"]
        assert sort["ms"] > 0 and sort["count_columns"] < hist["count_columns"]
        assert hist["ms"] > 0 and hist["count_columns"] == 2
        autotune.save(res, tmp_path / "tuning.json")
        autotune.apply({"hist_max_bins": 1024, "count_block": 512})
        assert autotune.current()["hist_max_bins"] == 1024
        assert autotune.load(tmp_path / "tuning.json") == res["best"]
        assert autotune.current() == res["best"]
        p = _capi.Plan(m, gid, T, opts, form=1)
        try:
            p.run(1)
            U1, P = p.results()
        finally:
            p.close()
    finally:
        m.close()
    o = oracle.mwu(data, rows, indptr, gid, T, n_rows=R, tie_correction=True,
                   use_continuity=True, zero_handling="min", alternative="two-sided",
                   method="asymptotic", with_stats=True)
    np.testing.assert_array_equal(U1, o[0])
    np.testing.assert_allclose(P, o[1], rtol=1e-12, atol=0)
