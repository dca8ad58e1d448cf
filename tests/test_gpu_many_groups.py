"""Jobs with thousands of groups (a perturbation screen: one target per
guide).  The reference keeps per-group state on the heap and takes any number
of groups (src/cpp/mannwhitneyu.cpp:855-874).  Here the sort path's split
keeps its group counters in LDS while they fit (up to about 12,900 groups)
and in HBM beyond; group means and the numba pipeline's sums take slices of
groups per workgroup.  U1 and tie bit-exact, p within 1e-12, means within
1e-6 of the oracle."""
import sys

import numpy as np
import pytest

import oracle
from fixtures import rel_close
from hpdex_amd import _capi, synthetic
from oracle import de as ode

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def screen(seed, R, T, dens, n_ref, f64=False, ints=False):
    """T targets of 3 to 6 cells each, a reference of n_ref cells, the other
    cells unlabelled; float values (or small integer counts) with ties."""
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.asarray(dens, float))
    if ints:
        data = rng.negative_binomial(1.0, 0.3, len(rows)) + 1.0
    else:
        data = np.round(rng.normal(0, 2, len(rows)), 2)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:n_ref]] = 0
    sizes = rng.integers(3, 7, T)
    lab = np.repeat(np.arange(1, T + 1, dtype=np.int32), sizes)
    assert n_ref + len(lab) <= R
    gid[perm[n_ref:n_ref + len(lab)]] = lab
    return dict(data=data.astype(np.float64 if f64 else np.float32), indices=rows,
                indptr=indptr, group_id=gid, n_targets=T, n_rows=R)


def check(w, mode):
    opts = _capi.make_opts(tie_correction=True, use_continuity=True,
                           zero_handling={"min": 1, "mix": 3}[mode])
    g = _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"], opts,
                      n_rows=w["n_rows"], with_stats=True)
    o = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], w["n_targets"],
                   n_rows=w["n_rows"], tie_correction=True, use_continuity=True,
                   zero_handling=mode, with_stats=True, threads=8)
    for a, b in zip((g[0], g[2], g[3], g[4]), (o[0], o[2], o[3], o[4])):
        assert np.array_equal(a, b)
    assert rel_close(g[1], o[1], rtol=1e-12)


@pytest.mark.parametrize("T", [8000, 14000], ids=["lds-counters", "hbm-counters"])
@pytest.mark.parametrize("f64", [False, True], ids=["f32", "f64"])
def test_sort_path_thousands_of_targets(T, f64):
    """8,000 targets (counters in the split's LDS, beyond the 64 KiB that
    used to bound them) and 14,000 (counters in HBM); a dense column (keys to
    fixed slots) and sparse ones (keys to offsets)."""
    say(f"many targets T={T} f64={f64}")
    w = screen(21 + T, 6 * T + 6000, T, [1.0, 0.5, 0.2], 4000, f64)
    check(w, "min")
    check(w, "mix")


@pytest.mark.parametrize("T", [8000, 14000], ids=["lds-counters", "hbm-counters"])
def test_histogram_branch_many_targets(T):
    """Integer counts on the sort path (count kernel off): the split turns a
    dense column's slot keys into bins and scatters sparse columns' bins."""
    say(f"many targets, histogram branch T={T}")
    t = _capi.get_tuning()
    saved = {n: int(getattr(t, n)) for n, _ in _capi.Tuning._fields_}
    _capi.set_tuning(count_lds_bytes=0)
    try:
        w = screen(25 + T, 6 * T + 6000, T, [1.0, 0.5, 0.2], 4000, ints=True)
        check(w, "min")
    finally:
        _capi.set_tuning(**saved)


def test_integer_counts_many_targets_count_path():
    """Integer counts with 14,000 targets: the count kernel takes them (up to
    65,535 groups)."""
    say("many targets, count path")
    w = screen(22, 90000, 14000, [0.6, 0.3, 0.1], 4000, ints=True)
    check(w, "min")


def test_group_mean_slices_of_groups():
    """10,001 groups: three slices of at most 4,096 groups per column."""
    say("group_mean many groups")
    w = screen(23, 70000, 10000, [0.9, 0.4, 0.05], 5000)
    G = w["n_targets"] + 1
    for inc in (True, False):
        got = _capi.group_mean_csc(w["data"], w["indices"], w["indptr"], w["group_id"], G,
                                   include_zeros=inc)
        want = oracle.group_mean(w["data"], w["indices"], w["indptr"], w["group_id"], G,
                                 include_zeros=inc)
        assert rel_close(got, want, rtol=1e-6, atol=1e-300)


def test_de_many_targets():
    """hpdex_de_csc with 6,000 targets: means, fold changes, U1, p and FDR."""
    say("de many targets")
    w = screen(24, 45000, 6000, [0.8, 0.3], 4000)
    T, C = w["n_targets"], len(w["indptr"]) - 1
    got = _capi.de_csc(w["data"], w["indices"], w["indptr"], w["group_id"], T,
                       n_rows=w["n_rows"])
    means = oracle.group_mean(w["data"], w["indices"], w["indptr"], w["group_id"], T + 1,
                              include_zeros=True)
    fc, lfc = ode.fold_change(means, C, T, 20.0)
    U1, P = oracle.mwu(w["data"], w["indices"], w["indptr"], w["group_id"], T,
                       n_rows=w["n_rows"], tie_correction=True, use_continuity=True,
                       zero_handling="min", threads=8)
    assert np.array_equal(got[0], U1)
    assert rel_close(got[1], P, rtol=1e-12)
    assert rel_close(got[2], fc, rtol=1e-12)
    assert rel_close(got[3], lfc, rtol=1e-12, atol=1e-15)
    assert rel_close(got[4], ode.fdr_bh(P), rtol=1e-12)
