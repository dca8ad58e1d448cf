"""A resident matrix's kept reference tables: after a job on the matrix, a
later job with the same reference rows (a new target set against the same
reference) ranks its sort-path columns from the kept sorted reference instead
of sorting it again.  Results must equal those of a job that sorts its own
reference bit for bit, and the oracle's (U1 and tie bit-exact, p within
1e-12); tables for other reference rows are never used."""
import sys

import numpy as np
import pytest

import oracle
from fixtures import rel_close
from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu

ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def keep_tuning():
    t = _capi.get_tuning()
    saved = {n: int(getattr(t, n)) for n, _ in _capi.Tuning._fields_}
    yield
    _capi.set_tuning(**saved)


def float_matrix(seed, R, dens, f64=False, ints=None):
    """Float columns of the given densities (values with ties), and integer
    columns (counts) when `ints` gives their densities."""
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.asarray(dens, float))
    data = np.round(rng.normal(0, 2, len(rows)), 2)
    if ints is not None:
        irows, iptr = synthetic._bernoulli_rows(rng, R, np.asarray(ints, float))
        ivals = rng.negative_binomial(0.8, 0.1, len(irows)) + 1.0
        rows = np.concatenate([rows, irows])
        indptr = np.concatenate([indptr, indptr[-1] + iptr[1:]])
        data = np.concatenate([data, ivals])
    return data.astype(np.float64 if f64 else np.float32), rows.astype(np.int64), indptr


def labels(rng, R, n_ref, T, ref=None, big=()):
    """Reference rows (`ref`, or n_ref drawn) and T targets over the other
    rows; targets in `big` take 300 rows each (ranked by the sorted pass)."""
    gid = np.full(R, -1, np.int32)
    if ref is None:
        ref = rng.choice(R, n_ref, replace=False)
    gid[ref] = 0
    rest = rng.permutation(np.nonzero(gid < 0)[0])
    k = 0
    for t in big:
        gid[rest[k:k + 300]] = t
        k += 300
    small = [t for t in range(1, T + 1) if t not in big]
    n = min(len(rest) - k, 40 * len(small))
    gid[rest[k:k + n]] = rng.choice(small, n)
    return gid, ref


def oracle_of(data, rows, indptr, gid, T, R, mode):
    return oracle.mwu(data, rows, indptr, gid, T, n_rows=R, tie_correction=True,
                      use_continuity=True, zero_handling=mode, with_stats=True, threads=8)


def run_plan(m, gid, T, mode, form=1):
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=ZH[mode])
    p = _capi.Plan(m, gid, T, opts, form=form, stats=True)
    info = p.info()
    p.run(2)
    out = p.results(with_stats=True)
    p.close()
    return out, info


def check_oracle(g, o):
    U1g, Pg, Tg, n1g, n2g = g
    U1o, Po, To, n1o, n2o = o
    assert np.array_equal(n1g, n1o) and np.array_equal(n2g, n2o)
    assert np.array_equal(U1g, U1o)
    assert np.array_equal(Tg, To)
    assert rel_close(Pg, Po, rtol=1e-12)


@pytest.mark.parametrize("f64", [False, True], ids=["f32", "f64"])
@pytest.mark.parametrize("mode", ["min", "mix", "none"])
def test_requery_uses_kept_reference(f64, mode, keep_tuning):
    """Reference slices on chip and beyond it (the cutover at 1,024 keys),
    targets of a few keys (windows) and of 300 (sorted pass): the follow-up
    job takes every sort-path column's tables from the matrix and matches a
    job that sorts again, and the oracle."""
    say(f"refcache requery f64={f64} mode={mode}")
    rng = np.random.default_rng(5 + f64)
    R, T = 9000, 24
    # references of ~2,500 values (sorted in HBM) and ~900 (on chip); at
    # these densities every target holds two values or more (mode none)
    data, rows, indptr = float_matrix(31, R, [0.95, 0.7, 0.5, 0.35], f64)
    C = len(indptr) - 1
    _capi.set_tuning(ref_cache=1, ref_lds_keys=1024)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    assert m.ref_cache_info() == {"columns": 0, "bytes": 0}
    gid_a, ref = labels(rng, R, 2600, T, big=(3,))
    ga, ia = run_plan(m, gid_a, T, mode)
    assert ia["split_columns"] == C and ia["ref_cached_columns"] == 0
    kept = m.ref_cache_info()
    assert kept["columns"] == C and kept["bytes"] > 0
    check_oracle(ga, oracle_of(data, rows, indptr, gid_a, T, R, mode))
    gid_b, _ = labels(rng, R, 0, T, ref=ref, big=(5, 7))
    for form in (1, 0):
        gb, ib = run_plan(m, gid_b, T, mode, form=form)
        assert ib["ref_cached_columns"] == C
        _capi.set_tuning(ref_cache=0)
        gs, is_ = run_plan(m, gid_b, T, mode, form=form)
        _capi.set_tuning(ref_cache=1)
        assert is_["ref_cached_columns"] == 0
        for a, b in zip(gb, gs):
            assert np.array_equal(a, b)
    check_oracle(gb, oracle_of(data, rows, indptr, gid_b, T, R, mode))
    m.close()


def test_other_reference_rows_are_not_served(keep_tuning):
    """A job with different reference rows sorts its own reference (and the
    matrix then keeps its tables instead); clearing drops them."""
    say("refcache other reference")
    rng = np.random.default_rng(8)
    R, T = 6000, 10
    data, rows, indptr = float_matrix(32, R, [0.8, 0.4, 0.05])
    _capi.set_tuning(ref_cache=1)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    gid_a, ref_a = labels(rng, R, 1500, T)
    run_plan(m, gid_a, T, "min")
    # one reference row less: not the same reference
    gid_c = gid_a.copy()
    gid_c[ref_a[0]] = -1
    gc, ic = run_plan(m, gid_c, T, "min")
    assert ic["ref_cached_columns"] == 0
    check_oracle(gc, oracle_of(data, rows, indptr, gid_c, T, R, "min"))
    # the matrix now keeps gid_c's reference: gid_a sorts again, gid_c does not
    assert run_plan(m, gid_a, T, "min")[1]["ref_cached_columns"] == 0
    assert run_plan(m, gid_a, T, "min")[1]["ref_cached_columns"] == 3
    m.ref_cache_clear()
    assert m.ref_cache_info() == {"columns": 0, "bytes": 0}
    assert run_plan(m, gid_a, T, "min")[1]["ref_cached_columns"] == 0
    m.close()


def test_histogram_branch_columns_are_not_kept(keep_tuning):
    """With the count kernel off, integer columns take the sort path's
    histogram branch, whose tables depend on the targets' values too: only the
    float columns are kept, and a re-query still matches the oracle."""
    say("refcache histogram branch")
    rng = np.random.default_rng(9)
    R, T = 7000, 12
    data, rows, indptr = float_matrix(33, R, [0.7, 0.2], ints=[0.5, 0.1, 0.3])
    _capi.set_tuning(ref_cache=1, count_lds_bytes=0)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    gid_a, ref = labels(rng, R, 1800, T)
    run_plan(m, gid_a, T, "mix")
    assert m.ref_cache_info()["columns"] == 2
    gid_b, _ = labels(rng, R, 0, T, ref=ref)
    gb, ib = run_plan(m, gid_b, T, "mix")
    assert ib["split_columns"] == 5 and ib["ref_cached_columns"] == 2
    check_oracle(gb, oracle_of(data, rows, indptr, gid_b, T, R, "mix"))
    m.close()


def test_failed_job_keeps_nothing(keep_tuning):
    """A job that fails (a target too small) leaves the matrix without tables."""
    say("refcache failed job")
    rng = np.random.default_rng(10)
    R, T = 4000, 6
    data, rows, indptr = float_matrix(34, R, [0.6, 0.3])
    _capi.set_tuning(ref_cache=1)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    gid, _ = labels(rng, R, 900, T)
    gid[gid == 4] = -1
    gid[np.nonzero(gid < 0)[0][0]] = 4   # target 4 holds one row
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=ZH["min"])
    p = _capi.Plan(m, gid, T, opts, form=1)
    with pytest.raises(_capi.HpdexError, match="Sample too small"):
        p.run(1)
        p.results()
    p.close()
    assert m.ref_cache_info()["columns"] == 0
    m.close()


def test_session_requery_uses_kept_reference(keep_tuning):
    """Session jobs on a resident matrix: the first keeps its reference, the
    follow-ups with the same reference rows use it; every job's results equal
    the same job run alone through a plan that sorts its own reference."""
    say("refcache session")
    rng = np.random.default_rng(12)
    R, T = 8000, 16
    data, rows, indptr = float_matrix(35, R, [0.9, 0.5, 0.2, 0.05])
    _capi.set_tuning(ref_cache=1)
    m = _capi.DeviceMatrix(data, rows, indptr, R)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=ZH["min"])
    gid_a, ref = labels(rng, R, 2000, T, big=(2,))
    s = _capi.Session(4 << 30, slots=2)
    U1a, Pa, _ = s.wait(s.submit(gid_a, T, opts, matrix=m))
    assert m.ref_cache_info()["columns"] == 4
    gids = [labels(rng, R, 0, T, ref=ref)[0] for _ in range(4)]
    tickets = [s.submit(g, T, opts, matrix=m) for g in gids]
    got = [s.wait(t)[:2] for t in tickets]
    s.close()
    _capi.set_tuning(ref_cache=0)
    for g, (U1, P) in zip([gid_a] + gids, [(U1a, Pa)] + got):
        alone = run_plan(m, g, T, "min")[0]
        assert np.array_equal(U1, alone[0]) and np.array_equal(P, alone[1])
    o = oracle_of(data, rows, indptr, gids[-1], T, R, "min")
    assert np.array_equal(got[-1][0], o[0]) and rel_close(got[-1][1], o[1], rtol=1e-12)
    m.close()
